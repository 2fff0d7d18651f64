// dwt_dev.h -- per-pair arithmetic of the periodic DWT (convolution
// alignment of reference src/la8.cpp:26-115 and the Haar/D4/C6
// equivalents), shared by the plain and batched kernels.
#pragma once

#include "wh_common.h"

namespace wh {

// One forward pair (s, d) from src[i0 .. i0+L-1] (indices already reduced).
template <typename T, int FILT>
__device__ __forceinline__ void fwd_pair(const T* v, T& s, T& d) {
  constexpr int L = Filter<FILT>::L;
  if constexpr (FILT == kHAAR) {
    s = (v[0] + v[1]) * T(kInvR2);
    d = (v[0] - v[1]) * T(kInvR2);
  } else {
    s = v[0] * T(Filter<FILT>::c(0));
    d = v[0] * T(gcoef<FILT>(0));
#pragma unroll
    for (int k = 1; k < L; ++k) {
      s = fma(v[k], T(Filter<FILT>::c(k)), s);
      d = fma(v[k], T(gcoef<FILT>(k)), d);
    }
  }
}

// One inverse pair from sv[q], dv[q] = s[m-q], d[m-q], q = 0..L/2-1.
template <typename T, int FILT>
__device__ __forceinline__ void inv_pair(const T* sv, const T* dv, T& e, T& o) {
  constexpr int H = Filter<FILT>::L / 2;
  if constexpr (FILT == kHAAR) {
    e = (sv[0] + dv[0]) * T(kInvR2);
    o = (sv[0] - dv[0]) * T(kInvR2);
  } else {
    e = sv[H - 1] * T(Filter<FILT>::c(2 * (H - 1)));
    o = sv[H - 1] * T(Filter<FILT>::c(2 * (H - 1) + 1));
    e = fma(dv[H - 1], T(gcoef<FILT>(2 * (H - 1))), e);
    o = fma(dv[H - 1], T(gcoef<FILT>(2 * (H - 1) + 1)), o);
#pragma unroll
    for (int q = H - 2; q >= 0; --q) {
      e = fma(sv[q], T(Filter<FILT>::c(2 * q)), e);
      o = fma(sv[q], T(Filter<FILT>::c(2 * q + 1)), o);
      e = fma(dv[q], T(gcoef<FILT>(2 * q)), e);
      o = fma(dv[q], T(gcoef<FILT>(2 * q + 1)), o);
    }
  }
}

}  // namespace wh
