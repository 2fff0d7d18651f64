// blocksel.h -- exact median / MAD of one signal's detail coefficients by a
// single workgroup (the batched pipeline smooths one signal per workgroup).
//
// MSB-first radix selection over order-preserving integer keys with 11-bit
// digits and an LDS histogram, tracking the two middle ranks (count-1)/2
// and count/2 at once; median = 0.5 (a + b) and MAD = 1.4826 median|v - med|
// as in reference src/utils.cpp:166-203.
#pragma once

#include "wh_common.h"

namespace wh {

constexpr int kSelBits = 11;
constexpr int kSelBins = 1 << kSelBits;

__device__ __forceinline__ uint64_t sel_key(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t sel_key(float v) {
  uint32_t u = __float_as_uint(v);
  return (u >> 31) ? uint64_t(~u) : uint64_t(u | 0x80000000u);
}
__device__ __forceinline__ double sel_val(uint64_t k, double*) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(int64_t(u));
}
__device__ __forceinline__ double sel_val(uint64_t k, float*) {
  uint32_t kk = uint32_t(k);
  uint32_t u = (kk >> 31) ? (kk & 0x7fffffffu) : ~kk;
  return double(__uint_as_float(u));
}

// LDS workspace of block_median (bytes)
template <int NT>
constexpr size_t block_sel_lds() {
  return size_t(2) * kSelBins * sizeof(unsigned) + size_t(NT) * sizeof(unsigned) + 64;
}

// Median of v[off + e * stride], e < count (|v - center| when deviation),
// computed by all NT threads of the workgroup; returns the same value in
// every thread.  ws: block_sel_lds<NT>() bytes of LDS.
template <typename T, int NT>
__device__ double block_median(const T* __restrict__ v, int64_t off, int64_t stride, int64_t count,
                               bool deviation, double center, unsigned char* ws) {
  unsigned* hist = reinterpret_cast<unsigned*>(ws);            // [2][kSelBins]
  unsigned* part = hist + 2 * kSelBins;                         // [NT]
  uint64_t* st = reinterpret_cast<uint64_t*>(part + NT);        // prefix0, prefix1, found0, found1
  long long* ks = reinterpret_cast<long long*>(st + 4);         // k0, k1
  const int tid = threadIdx.x;
  const T c = T(center);
  if (tid == 0) {
    st[0] = st[1] = 0;
    ks[0] = (count - 1) / 2;
    ks[1] = count / 2;
  }
  constexpr int kbits = int(sizeof(T)) * 8;
  for (int hi = kbits; hi > 0;) {
    const int lo = hi > kSelBits ? hi - kSelBits : 0;
    for (int i = tid; i < 2 * kSelBins; i += NT) hist[i] = 0;
    __syncthreads();
    const uint64_t p0 = st[0], p1 = st[1];
    const bool same = (p0 == p1);
    const uint64_t dmask = (uint64_t(1) << (hi - lo)) - 1;
    for (int64_t e = tid; e < count; e += NT) {
      T x = v[off + e * stride];
      if (deviation) x = fabs(x - c);
      const uint64_t k = sel_key(x);
      const uint64_t top = (hi >= 64) ? 0 : (k >> hi);
      const unsigned d = unsigned((k >> lo) & dmask);
      if (top == p0) atomicAdd(&hist[d], 1u);
      if (!same && top == p1) atomicAdd(&hist[kSelBins + d], 1u);
    }
    __syncthreads();
    constexpr int per = kSelBins / NT;
    for (int r = 0; r < 2; ++r) {
      const unsigned* hr = hist + (same ? 0 : r) * kSelBins;
      const long long k = ks[r];  // read before the barriers; the finder updates it below
      unsigned loc = 0;
#pragma unroll
      for (int i = 0; i < per; ++i) loc += hr[tid * per + i];
      part[tid] = loc;
      __syncthreads();
      for (int o = 1; o < NT; o <<= 1) {
        unsigned add = (tid >= o) ? part[tid - o] : 0u;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
      }
      const long long before = (tid == 0) ? 0 : (long long)part[tid - 1];
      if (before <= k && k < (long long)part[tid]) {
        long long acc = before;
        for (int i = 0; i < per; ++i) {
          const unsigned cnt = hr[tid * per + i];
          if (acc + cnt > k) {
            st[2 + r] = uint64_t(tid * per + i);
            ks[r] = k - acc;
            break;
          }
          acc += cnt;
        }
      }
      __syncthreads();
      if (tid == 0) st[r] = (st[r] << (hi - lo)) | st[2 + r];
      __syncthreads();
    }
    hi = lo;
  }
  const double a = sel_val(st[0], (T*)nullptr), b = sel_val(st[1], (T*)nullptr);
  __syncthreads();
  return 0.5 * (a + b);
}

// 1.4826 median|v - median(v)|
template <typename T, int NT>
__device__ double block_mad(const T* __restrict__ v, int64_t off, int64_t stride, int64_t count,
                            unsigned char* ws) {
  const double med = block_median<T, NT>(v, off, stride, count, false, 0.0, ws);
  return 1.4826 * block_median<T, NT>(v, off, stride, count, true, med, ws);
}

}  // namespace wh
