// Instantiation unit of the fused MODWT kernels: element type float, filter D4.
#include "modwt_launch.h"

namespace wh {
template <>
int modwt_launch<float, kD4>(bool inverse, const GroupLaunch& gl, const ModwtArgs& a,
                            hipStream_t st) {
  return modwt_launch_impl<float, kD4>(inverse, gl, a, st);
}
}  // namespace wh
