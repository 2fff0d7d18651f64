// batch_launch.h -- launch of the persistent batch kernel for one element
// type; included by the batch_<type>.hip instantiation units only.
#pragma once

#include <mutex>

#include "batch_kern.h"
#include "wh_internal.h"

namespace wh {

template <typename T, int NT>
int batch_launch_nt(const BatchLaunch& bl, const BatchArgs& a, hipStream_t st) {
  auto kern = batch_smooth_k<T, NT>;
  allow_dynamic_lds(reinterpret_cast<const void*>(kern));
  hipLaunchKernelGGL(kern, dim3(bl.grid), dim3(NT), bl.lds, st, a);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

template <typename T, int NT>
int batch_occ_nt(size_t lds, int* blocks) {
  auto kern = batch_smooth_k<T, NT>;
  allow_dynamic_lds(reinterpret_cast<const void*>(kern));
  WH_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kern, NT, lds));
  return WH_OK;
}

template <typename T>
int batch_occupancy_impl(int NT, size_t lds, int* blocks) {
  switch (NT) {
    case 128: return batch_occ_nt<T, 128>(lds, blocks);
    case 256: return batch_occ_nt<T, 256>(lds, blocks);
    case 512: return batch_occ_nt<T, 512>(lds, blocks);
    default: return set_error(WH_ERR_UNSUPPORTED, "batch: unsupported workgroup size");
  }
}

template <typename T>
int batch_launch_impl(const BatchLaunch& bl, const BatchArgs& a, hipStream_t st) {
  switch (bl.NT) {
    case 128: return batch_launch_nt<T, 128>(bl, a, st);
    case 256: return batch_launch_nt<T, 256>(bl, a, st);
    case 512: return batch_launch_nt<T, 512>(bl, a, st);
    default: return set_error(WH_ERR_UNSUPPORTED, "batch: unsupported workgroup size");
  }
}

}  // namespace wh
