// Instantiation unit of the fused MODWT kernels: element type float, filter HAAR.
#include "modwt_launch.h"

namespace wh {
template <>
int modwt_launch<float, kHAAR>(bool inverse, const GroupLaunch& gl, const ModwtArgs& a,
                            hipStream_t st) {
  return modwt_launch_impl<float, kHAAR>(inverse, gl, a, st);
}
}  // namespace wh
