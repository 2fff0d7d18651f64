// Instantiation unit of the fused MODWT kernels: element type double, filter C6.
#include "modwt_launch.h"

namespace wh {
template <>
int modwt_launch<double, kC6>(bool inverse, const GroupLaunch& gl, const ModwtArgs& a,
                            hipStream_t st) {
  return modwt_launch_impl<double, kC6>(inverse, gl, a, st);
}
}  // namespace wh
