// Instantiation unit of the persistent batch kernel: element type float.
#include "batch_launch.h"

namespace wh {
template <>
int batch_launch<float>(const BatchLaunch& bl, const BatchArgs& a, hipStream_t st) {
  return batch_launch_impl<float>(bl, a, st);
}
template <>
int batch_occupancy<float>(int NT, size_t lds, int* blocks) {
  return batch_occupancy_impl<float>(NT, lds, blocks);
}
}  // namespace wh
