// Instantiation unit of the fused MODWT kernels: element type double, filter LA8.
#include "modwt_launch.h"

namespace wh {
template <>
int modwt_launch<double, kLA8>(bool inverse, const GroupLaunch& gl, const ModwtArgs& a,
                            hipStream_t st) {
  return modwt_launch_impl<double, kLA8>(inverse, gl, a, st);
}
}  // namespace wh
