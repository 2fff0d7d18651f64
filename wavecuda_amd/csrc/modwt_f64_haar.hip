// Instantiation unit of the fused MODWT kernels: element type double, filter HAAR.
#include "modwt_launch.h"

namespace wh {
template <>
int modwt_launch<double, kHAAR>(bool inverse, const GroupLaunch& gl, const ModwtArgs& a,
                            hipStream_t st) {
  return modwt_launch_impl<double, kHAAR>(inverse, gl, a, st);
}
}  // namespace wh
