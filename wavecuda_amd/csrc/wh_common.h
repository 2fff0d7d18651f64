// wh_common.h -- shared definitions of the wavehip HIP library (gfx950).
//
// Filter tables, codes and small device helpers used by every kernel file.
// Codes follow the reference's R glue (src/wvtheads.h:14-38,
// src/rfunctions.cuh:13-15, R/rfunctions.R:7-37).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wavehip.h"

namespace wh {

enum { kFWD = 1, kBWD = 0 };
enum { kDWT = 0, kMODWT = 1, kMODWT_PO = 2 };  // kMODWT: time ordered
enum { kHAAR = 1, kD4 = 2, kC6 = 3, kLA8 = 4 };
enum { kHARD = 0, kSOFT = 1 };

// Scaling filters with the reference's digits (src/haarcoeffs.h:4-5,
// src/daub4coeffs.h:6-9, src/c6coeffs.h:6-11, src/la8coeffs.h:6-13);
// the wavelet filter is g_k = (-1)^k C_{L-1-k} (src/daub4.cpp:36).
template <int FILT>
struct Filter;

template <>
struct Filter<kHAAR> {
  static constexpr int L = 2;
  __host__ __device__ static constexpr double c(int k) {
    return k == 0 ? 0.70710678118654752440 : 0.70710678118654752440;
  }
};
template <>
struct Filter<kD4> {
  static constexpr int L = 4;
  __host__ __device__ static constexpr double c(int k) {
    return k == 0   ? 0.4829629131445341
           : k == 1 ? 0.8365163037378079
           : k == 2 ? 0.2241438680420134
                    : -0.1294095225512604;
  }
};
template <>
struct Filter<kC6> {
  static constexpr int L = 6;
  __host__ __device__ static constexpr double c(int k) {
    return k == 0   ? -7.273261951252645e-2
           : k == 1 ? 3.378976624574818e-1
           : k == 2 ? 8.525720202116004e-1
           : k == 3 ? 3.848648468648577e-1
           : k == 4 ? -7.273261951252645e-2
                    : -1.565572813579199e-2;
  }
};
template <>
struct Filter<kLA8> {
  static constexpr int L = 8;
  __host__ __device__ static constexpr double c(int k) {
    return k == 0   ? -7.576571478950221e-2
           : k == 1 ? -2.963552764600249e-2
           : k == 2 ? 4.976186676327750e-1
           : k == 3 ? 8.037387518051321e-1
           : k == 4 ? 2.978577956053061e-1
           : k == 5 ? -9.921954357663353e-2
           : k == 6 ? -1.260396726203130e-2
                    : 3.222310060405147e-2;
  }
};

template <int FILT>
__host__ __device__ constexpr double gcoef(int k) {
  return (k & 1) ? -Filter<FILT>::c(Filter<FILT>::L - 1 - k) : Filter<FILT>::c(Filter<FILT>::L - 1 - k);
}

constexpr double kInvR2 = 0.70710678118654752440;
constexpr double kHInvR2 = 0.35355339059327376220;

inline int filter_length(int filt) {
  switch (filt) {
    case kHAAR: return 2;
    case kD4: return 4;
    case kC6: return 6;
    case kLA8: return 8;
    default: return 0;
  }
}

// src/wavutils.cpp:48-65 (maxlevels = floor(log2(n/L)) + 1, 0 -> max).
inline int resolve_levels(int64_t n, int nlevels, int L) {
  if (n <= 0 || (n & (n - 1)) != 0 || L <= 0) return 0;
  int lg = 0;
  while ((int64_t(1) << lg) < n) ++lg;
  // floor(log2(n/L)) for power-of-two n: log2(n) - ceil(log2(L)) when L is
  // not a power of two, log2(n) - log2(L) when it is.
  int cl = 0;
  while ((1 << cl) < L) ++cl;
  int maxlev = lg - cl + 1;
  if (n < L) maxlev = 0;
  if (maxlev <= 0) return 0;
  if (nlevels == 0) return maxlev;
  if (nlevels < 0 || nlevels > maxlev) return 0;
  return nlevels;
}

// two consecutive elements as one 16-byte (double) or 8-byte (float) access
template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

// soft/hard rule of src/thresh.cpp:142-157
template <typename T>
__device__ __forceinline__ T thresh_coef(T c, T t, int hardness) {
  T a = fabs(c);
  if (a < t) return T(0);
  if (hardness == kSOFT) return (c < T(0)) ? -(a - t) : (a - t);
  return c;
}

__device__ __forceinline__ int64_t pmod(int64_t a, int64_t n) {
  // n is a power of two
  return a & (n - 1);
}

// Order-preserving integer keys of floating-point values (the radix order
// of the exact selections): negative values have all bits flipped,
// non-negative ones the sign bit set.  The first radix digit of a key is
// its top kSelBits = 11 bits (key >> 53 for FP64, >> 21 for FP32).
__device__ __forceinline__ uint64_t sel_key(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t sel_key(float v) {
  uint32_t u = __float_as_uint(v);
  return (u >> 31) ? uint64_t(~u) : uint64_t(u | 0x80000000u);
}
template <typename T>
__device__ __forceinline__ unsigned sel_digit0(T v) {
  return unsigned(sel_key(v) >> (sizeof(T) * 8 - 11));
}

}  // namespace wh

#define WH_HIP_CHECK(expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return wh::set_hip_error(_e, #expr); \
  } while (0)

namespace wh {
int set_hip_error(hipError_t e, const char* what);
int set_error(int code, const char* msg);
}  // namespace wh
