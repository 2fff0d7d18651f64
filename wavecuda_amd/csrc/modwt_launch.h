// modwt_launch.h -- kernel selection for one (element type, filter); included
// by the modwt_<type>_<filter>.hip instantiation units only.
#pragma once

#include <mutex>

#include "modwt_kern.h"
#include "wh_internal.h"

namespace wh {

template <typename KERN>
inline void modwt_allow_lds(KERN kern) {
  allow_dynamic_lds(reinterpret_cast<const void*>(kern));
}

template <typename T, int FILT, int R, int NT, int G>
int modwt_launch_one(bool inverse, const GroupLaunch& gl, const ModwtArgs& a, hipStream_t st) {
  const dim3 grid(unsigned(gl.blocks), unsigned(gl.count));
  if (inverse) {
    constexpr int MAP = (R == 1) ? 2 : 0;
    auto kern = modwt_inv_k<T, FILT, R, NT, G, MAP>;
    modwt_allow_lds(kern);
    hipLaunchKernelGGL(kern, grid, dim3(NT), gl.lds, st, a);
  } else if (gl.sp == kStoreAdjoint) {
    auto kern = modwt_fwd_k<T, FILT, R, NT, G, 0, kStoreAdjoint>;
    modwt_allow_lds(kern);
    hipLaunchKernelGGL(kern, grid, dim3(NT), gl.lds, st, a);
  } else {
    auto kern = modwt_fwd_k<T, FILT, R, NT, G, 0, kStoreAll>;
    modwt_allow_lds(kern);
    hipLaunchKernelGGL(kern, grid, dim3(NT), gl.lds, st, a);
  }
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

template <typename T, int FILT, int R, int NT, int G>
int modwt_launch_g(bool inverse, const GroupLaunch& gl, const ModwtArgs& a, hipStream_t st) {
  if constexpr (G > kGmax) {
    return set_error(WH_ERR_UNSUPPORTED, "modwt: group depth out of range");
  } else {
    if (gl.G == G) return modwt_launch_one<T, FILT, R, NT, G>(inverse, gl, a, st);
    return modwt_launch_g<T, FILT, R, NT, G + 1>(inverse, gl, a, st);
  }
}

// first inverse group with 256 threads (two workgroups per CU)
template <typename T, int FILT, int G>
int modwt_launch_inv256(const GroupLaunch& gl, const ModwtArgs& a, hipStream_t st) {
  if constexpr (G > kGmax) {
    return set_error(WH_ERR_UNSUPPORTED, "modwt: group depth out of range");
  } else {
    if (gl.G != G) return modwt_launch_inv256<T, FILT, G + 1>(gl, a, st);
    auto kern = modwt_inv_k<T, FILT, 1, 256, G, 2>;
    modwt_allow_lds(kern);
    hipLaunchKernelGGL(kern, dim3(unsigned(gl.blocks), unsigned(gl.count)), dim3(256), gl.lds, st,
                       a);
    WH_HIP_CHECK(hipGetLastError());
    return WH_OK;
  }
}

template <typename T, int FILT>
int modwt_launch_impl(bool inverse, const GroupLaunch& gl, const ModwtArgs& a, hipStream_t st) {
  if (gl.R == 1 && gl.NT == 512) return modwt_launch_g<T, FILT, 1, 512, 1>(inverse, gl, a, st);
  if (inverse && gl.R == 1 && gl.NT == 256) return modwt_launch_inv256<T, FILT, 1>(gl, a, st);
  if (gl.R == 32 && gl.NT == 256) return modwt_launch_g<T, FILT, 32, 256, 1>(inverse, gl, a, st);
  return set_error(WH_ERR_UNSUPPORTED, "modwt: unsupported residue tiling");
}

}  // namespace wh
