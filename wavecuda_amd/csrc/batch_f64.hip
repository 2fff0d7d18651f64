// Instantiation unit of the persistent batch kernel: element type double.
#include "batch_launch.h"

namespace wh {
template <>
int batch_launch<double>(const BatchLaunch& bl, const BatchArgs& a, hipStream_t st) {
  return batch_launch_impl<double>(bl, a, st);
}
template <>
int batch_occupancy<double>(int NT, size_t lds, int* blocks) {
  return batch_occupancy_impl<double>(NT, lds, blocks);
}
}  // namespace wh
