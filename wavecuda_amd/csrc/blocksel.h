// blocksel.h -- exact median / MAD of one signal's detail coefficients by a
// single workgroup (the batched pipeline smooths one signal per workgroup).
//
// MSB-first radix selection over order-preserving integer keys with 11-bit
// digits and an LDS histogram, tracking the two middle ranks (count-1)/2
// and count/2 at once; median = 0.5 (a + b) and MAD = 1.4826 median|v - med|
// as in reference src/utils.cpp:166-203.
//
// The values live in global memory (the workgroup's scratch).  Each pass
// over them issues kSelUnroll independent loads per thread before using
// any, so a pass costs a few memory latencies rather than one per element.
// As soon as the histogram shows that the values still matching the
// selected prefixes fit in LDS (kSelCap keys), one more pass copies their
// keys there and the remaining digits are resolved in LDS -- or, when at
// most kSelDirect keys are left (the usual case: a middle rank's bucket
// holds a few dozen values after one or two digits), every candidate's rank
// within its bucket is counted directly, one pass and one barrier instead
// of a radix digit's five barriers for each remaining digit.
//
// Two shortcuts, both exact: the first digit's counts of the median's values
// may come precomputed (the batched forward counts them as it produces the
// level-1 details), and the deviations |v - med|, whose keys crowd a few
// exponents, are first counted into equal-width bins over [0, bound], the
// bound following from the extreme first digits the values have.
#pragma once

#include "wh_common.h"

namespace wh {

constexpr int kSelBits = 11;
constexpr int kSelBins = 1 << kSelBits;
constexpr int kSelCap = 2048;   // keys kept in LDS
constexpr int kSelUnroll = 8;   // loads in flight per thread and pass
constexpr int kSelDirect = 256; // candidates ranked by direct counting

__device__ __forceinline__ double sel_val(uint64_t k, double*) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(int64_t(u));
}
__device__ __forceinline__ double sel_val(uint64_t k, float*) {
  uint32_t kk = uint32_t(k);
  uint32_t u = (kk >> 31) ? (kk & 0x7fffffffu) : ~kk;
  return double(__uint_as_float(u));
}

// bin of a value y in [0, hi] among kSelBins equal-width bins (scale =
// kSelBins / hi); monotone non-decreasing in y, clamped to the last bin
__device__ __forceinline__ int sel_bin(double y, double scale) {
  const double u = y * scale;
  return u > 0.0 ? (u < double(kSelBins - 1) ? int(u) : kSelBins - 1) : 0;
}

// smallest and largest digit with a non-zero count in h[0..kSelBins), by
// all NT threads (the same in every thread; -1 and kSelBins when empty).
// wk: 8 (NT / 64) bytes of LDS scratch.
template <int NT>
__device__ __forceinline__ void sel_digit_range(const unsigned* h, int* wk, int* dmin, int* dmax) {
  constexpr int per = kSelBins / NT;
  constexpr int kWaves = NT / 64;
  int lo = kSelBins, hi = -1;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    const int d = int(threadIdx.x) * per + i;
    if (h[d]) {
      lo = min(lo, d);
      hi = max(hi, d);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    wk[wave] = lo;
    wk[kWaves + wave] = hi;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    lo = min(lo, wk[w]);
    hi = max(hi, wk[kWaves + w]);
  }
  __syncthreads();  // wk is reused by the caller
  *dmin = lo;
  *dmax = hi;
}

// LDS layout of block_median: the candidate keys (8-byte aligned) first,
// then the two histograms, the scan array (the wave totals of the bucket
// scan) and the selection state (prefixes, ranks, counts, the copy
// counter, the values' first-digit range).
template <int NT>
constexpr size_t block_sel_lds() {
  return size_t(kSelCap) * 8 + size_t(2) * kSelBins * 4 + size_t(NT) * 4 + 128;
}
// where block_median leaves the smallest and largest first digit (2 ints)
template <int NT>
__device__ __forceinline__ int* sel_range_slot(unsigned char* ws) {
  return reinterpret_cast<int*>(ws + size_t(kSelCap) * 8 + size_t(2) * kSelBins * 4 +
                                size_t(NT) * 4 + 96);
}

// f(key) for every value v[off + e * stride], e < count (|v - center| when
// deviation).  Contiguous values (stride 1, v + off aligned to 16 bytes)
// are read 16 bytes at a time (two FP64 or four FP32 values per load);
// kSelUnroll loads in flight per thread either way.
template <typename T, int NT, typename F>
__device__ __forceinline__ void sel_for_each(const T* __restrict__ v, int64_t off, int64_t stride,
                                             int64_t count, bool deviation, T c, F f) {
  constexpr int kW = 16 / int(sizeof(T));
  typedef T TW __attribute__((ext_vector_type(kW)));
  auto key_of = [&](T y) { return sel_key(deviation ? T(fabs(y - c)) : y); };
  if (stride == 1 && reinterpret_cast<uintptr_t>(v + off) % sizeof(TW) == 0) {
    const TW* __restrict__ vw = reinterpret_cast<const TW*>(v + off);
    const int64_t nv = count / kW;
    for (int64_t b = threadIdx.x; b < nv; b += int64_t(NT) * kSelUnroll) {
      TW x[kSelUnroll];
#pragma unroll
      for (int u = 0; u < kSelUnroll; ++u) {
        const int64_t p = b + int64_t(u) * NT;
        x[u] = vw[p < nv ? p : b];
      }
#pragma unroll
      for (int u = 0; u < kSelUnroll; ++u) {
        const int64_t p = b + int64_t(u) * NT;
        if (p < nv) {
#pragma unroll
          for (int w = 0; w < kW; ++w) f(key_of(x[u][w]));
        }
      }
    }
    const int64_t tail = nv * kW;
    if (threadIdx.x < count - tail) f(key_of(v[off + tail + threadIdx.x]));
    return;
  }
  for (int64_t b = threadIdx.x; b < count; b += int64_t(NT) * kSelUnroll) {
    T x[kSelUnroll];
#pragma unroll
    for (int u = 0; u < kSelUnroll; ++u) {
      const int64_t e = b + int64_t(u) * NT;
      x[u] = (e < count) ? v[off + e * stride] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kSelUnroll; ++u) {
      const int64_t e = b + int64_t(u) * NT;
      if (e < count) f(key_of(x[u]));
    }
  }
}

// Median of v[off + e * stride], e < count (|v - center| when deviation),
// computed by all NT threads of the workgroup; returns the same value in
// every thread.  ws: block_sel_lds<NT>() bytes of LDS.  pre: optional
// counts of the values' first digits (sel_digit0), kSelBins of them, which
// then replace the first counting pass.  bound > 0 (deviations only): every
// value lies in [0, bound]; one pass then counts the values into kSelBins
// equal-width bins, and when the middle ranks' bins hold at most
// kSelDirect values they are copied to LDS and ranked directly (otherwise
// the radix digits start from the top).  With kRange, the smallest and
// largest first digit any value has are left at sel_range_slot(ws).
template <typename T, int NT, bool kRange = false>
__device__ double block_median(const T* __restrict__ v, int64_t off, int64_t stride, int64_t count,
                               bool deviation, double center, unsigned char* ws,
                               const unsigned* pre, double bound) {
  uint64_t* cand = reinterpret_cast<uint64_t*>(ws);                     // [kSelCap]
  unsigned* hist = reinterpret_cast<unsigned*>(cand + kSelCap);        // [2][kSelBins]
  unsigned* part = hist + 2 * kSelBins;                                 // [NT]
  uint64_t* st = reinterpret_cast<uint64_t*>(part + NT);  // prefix0, prefix1 (+2 spare)
  long long* ks = reinterpret_cast<long long*>(st + 4);   // k0, k1
  long long* cs = ks + 2;                                  // keys under the new prefixes
  int* ncand = reinterpret_cast<int*>(cs + 2);
  const int tid = threadIdx.x;
  const T c = T(center);
  if (tid == 0) {
    st[0] = st[1] = 0;
    ks[0] = (count - 1) / 2;
    ks[1] = count / 2;
    *ncand = 0;
  }
  constexpr int kbits = int(sizeof(T)) * 8;
  constexpr int per = kSelBins / NT;
  constexpr int kWaves = NT / 64;
  bool in_lds = false;
  int nc = 0;
  if (bound > 0.0) {
    // equal-width bins over [0, bound] (monotone in the value): the middle
    // ranks fall in bins b0 <= b1, with no value between them
    const double scale = double(kSelBins) / bound;
    for (int i = tid; i < kSelBins; i += NT) hist[i] = 0;
    __syncthreads();
    sel_for_each<T, NT>(v, off, stride, count, deviation, c, [&](uint64_t k) {
      atomicAdd(&hist[sel_bin(sel_val(k, (T*)nullptr), scale)], 1u);
    });
    __syncthreads();
    uint64_t* wsum = reinterpret_cast<uint64_t*>(part);  // [kWaves]
    const long long k0 = ks[0], k1 = ks[1];
    uint64_t loc = 0;
#pragma unroll
    for (int i = 0; i < per; ++i) loc += hist[tid * per + i];
    uint64_t inc = loc;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint64_t base = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
      if (w < wave) base += wsum[w];
    const long long before = (long long)(base + inc - loc);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const long long k = r ? k1 : k0;
      if (before <= k && k < before + (long long)loc) {
        long long acc = before;
        for (int i = 0; i < per; ++i) {
          const unsigned cnt = hist[tid * per + i];
          if (acc + cnt > k) {
            st[2 + r] = uint64_t(tid * per + i);
            ks[r] = k - acc;
            cs[r] = cnt;
            break;
          }
          acc += cnt;
        }
      }
    }
    __syncthreads();
    const int b0 = int(st[2]), b1 = int(st[3]);
    const long long total = cs[0] + (b1 != b0 ? cs[1] : 0);
    if (total <= kSelDirect) {
      const long long r0 = ks[0], r1 = (b1 != b0 ? cs[0] : 0) + ks[1];
      sel_for_each<T, NT>(v, off, stride, count, deviation, c, [&](uint64_t k) {
        const int q = sel_bin(sel_val(k, (T*)nullptr), scale);
        if (q >= b0 && q <= b1) cand[atomicAdd(ncand, 1)] = k;
      });
      __syncthreads();
      nc = *ncand;
      in_lds = true;
      // every candidate is ranked among all of them (prefix 0 of 0 bits)
      if (tid == 0) {
        ks[0] = r0;
        ks[1] = r1;
      }
    } else if (tid == 0) {  // bins too full: radix digits from the top
      ks[0] = (count - 1) / 2;
      ks[1] = count / 2;
    }
    __syncthreads();
  }
  for (int hi = kbits; hi > 0;) {
    if (in_lds && nc <= kSelDirect) {
      // rank of each candidate among the candidates of its bucket: the key
      // of rank k in bucket q is the x with lt <= k < lt + eq
      const uint64_t q0 = st[0], q1 = st[1];
      const long long k0 = ks[0], k1 = ks[1];
      __syncthreads();  // every thread has read st and ks before they change
      // counts of candidate x among the candidates j in [j0, j1)
      auto count = [&](uint64_t x, int j0, int j1, int* c) {
        c[0] = c[1] = c[2] = c[3] = 0;
        for (int j = j0; j < j1; ++j) {
          const uint64_t y = cand[j], ty = hi >= 64 ? 0 : y >> hi;
          if (ty == q0) {
            c[0] += (y < x);
            c[1] += (y == x);
          }
          if (ty == q1) {
            c[2] += (y < x);
            c[3] += (y == x);
          }
        }
      };
      auto decide = [&](uint64_t x, const int* c) {
        const uint64_t tx = hi >= 64 ? 0 : x >> hi;
        if (tx == q0 && c[0] <= k0 && k0 < c[0] + c[1]) st[0] = x;
        if (tx == q1 && c[2] <= k1 && k1 < c[2] + c[3]) st[1] = x;
      };
      if (2 * nc <= NT) {
        // two threads per candidate, each counting half of the others
        // (partial counts in the histograms' space, unused here)
        int* pc = reinterpret_cast<int*>(hist);  // [4][NT]
        const int half = tid / (NT / 2), i = tid % (NT / 2);
        const int jm = nc / 2;
        if (i < nc) {
          int c[4];
          count(cand[i], half ? jm : 0, half ? nc : jm, c);
#pragma unroll
          for (int q = 0; q < 4; ++q) pc[q * NT + tid] = c[q];
        }
        __syncthreads();
        if (tid < nc) {
          int c[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) c[q] = pc[q * NT + tid] + pc[q * NT + tid + NT / 2];
          decide(cand[tid], c);
        }
      } else {
        for (int i = tid; i < nc; i += NT) {
          int c[4];
          count(cand[i], 0, nc, c);
          decide(cand[i], c);
        }
      }
      __syncthreads();
      break;
    }
    const int lo = hi > kSelBits ? hi - kSelBits : 0;
    const bool counted = pre && hi == kbits;  // the first digit's counts are given
    if (!counted) {
      for (int i = tid; i < 2 * kSelBins; i += NT) hist[i] = 0;
    }
    __syncthreads();
    const uint64_t p0 = st[0], p1 = st[1];
    const bool same = (p0 == p1);
    const uint64_t dmask = (uint64_t(1) << (hi - lo)) - 1;
    auto count_key = [&](uint64_t k) {
      const uint64_t top = (hi >= 64) ? 0 : (k >> hi);
      const unsigned d = unsigned((k >> lo) & dmask);
      if (top == p0) atomicAdd(&hist[d], 1u);
      if (!same && top == p1) atomicAdd(&hist[kSelBins + d], 1u);
    };
    if (counted) {
      // nothing to count
    } else if (in_lds) {
      for (int i = tid; i < nc; i += NT) count_key(cand[i]);
    } else {
      sel_for_each<T, NT>(v, off, stride, count, deviation, c, count_key);
    }
    __syncthreads();
    // Both ranks' bucket counts packed in one word (rank 0 in the low 32
    // bits, rank 1 in the high; each is at most count < 2^32, so the low
    // half never carries), so one scan serves both: each thread sums its
    // `per` buckets, a wave scans by shuffles, then the wave totals.
    uint64_t* wsum = reinterpret_cast<uint64_t*>(part);  // [kWaves]
    const long long k0 = ks[0], k1 = ks[1];  // the finder below updates ks
    const unsigned* h0 = counted ? pre : hist;
    const unsigned* h1 = same ? h0 : hist + kSelBins;
    uint64_t loc = 0;
#pragma unroll
    for (int i = 0; i < per; ++i)
      loc += uint64_t(h0[tid * per + i]) | (uint64_t(h1[tid * per + i]) << 32);
    uint64_t inc = loc;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint64_t base = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
      if (w < wave) base += wsum[w];
    const uint64_t before = base + inc - loc;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const unsigned* hr = r ? h1 : h0;
      const long long k = r ? k1 : k0;
      const long long bef = (long long)(r ? (before >> 32) : (before & 0xffffffffull));
      const long long mine = (long long)(r ? (loc >> 32) : (loc & 0xffffffffull));
      if (bef <= k && k < bef + mine) {
        long long acc = bef;
        for (int i = 0; i < per; ++i) {
          const unsigned cnt = hr[tid * per + i];
          if (acc + cnt > k) {
            // the rank's prefix grows by this digit (every thread read the
            // old prefixes before counting)
            st[r] = ((r ? p1 : p0) << (hi - lo)) | uint64_t(tid * per + i);
            ks[r] = k - acc;
            cs[r] = cnt;
            break;
          }
          acc += cnt;
        }
      }
    }
    __syncthreads();
    // the first digit's counts cover every value: their extreme digits
    // bound the values (the copy area is free until a copy pass)
    if (kRange && hi == kbits) {
      int dmin, dmax;
      sel_digit_range<NT>(h0, reinterpret_cast<int*>(cand), &dmin, &dmax);
      if (tid == 0) {
        int* slot = sel_range_slot<NT>(ws);
        slot[0] = dmin;
        slot[1] = dmax;
      }
    }
    hi = lo;
    if (!in_lds && hi > 0) {
      const uint64_t q0 = st[0], q1 = st[1];
      const long long total = cs[0] + (q0 != q1 ? cs[1] : 0);
      if (total <= kSelCap) {
        // one more pass keeps the keys under the selected prefixes in LDS
        sel_for_each<T, NT>(v, off, stride, count, deviation, c, [&](uint64_t k) {
          const uint64_t top = k >> hi;
          if (top == q0 || top == q1) cand[atomicAdd(ncand, 1)] = k;
        });
        __syncthreads();
        nc = *ncand;
        in_lds = true;
      }
    }
  }
  const double a = sel_val(st[0], (T*)nullptr), b = sel_val(st[1], (T*)nullptr);
  __syncthreads();
  return 0.5 * (a + b);
}

// 1.4826 median|v - median(v)|
// pre: optional first-digit counts of the values (see block_median).  The
// smallest and largest first digits present (from pre or from the median's
// own first digit) bound the values, so the deviations' median can start
// from equal-width bins.
template <typename T, int NT>
__device__ double block_mad(const T* __restrict__ v, int64_t off, int64_t stride, int64_t count,
                            unsigned char* ws, const unsigned* pre = nullptr) {
  const bool binned = count > kSelDirect;
  // without pre the median's first digit gives the range
  const double med =
      (binned && !pre) ? block_median<T, NT, true>(v, off, stride, count, false, 0.0, ws, pre, 0.0)
                       : block_median<T, NT>(v, off, stride, count, false, 0.0, ws, pre, 0.0);
  int dr[2];
  if (binned && pre) {
    sel_digit_range<NT>(pre, reinterpret_cast<int*>(ws), &dr[0], &dr[1]);
  } else {
    const int* slot = sel_range_slot<NT>(ws);  // written before block_median's last barrier
    dr[0] = slot[0];
    dr[1] = slot[1];
  }
  double bound = 0.0;
  if (binned && dr[0] <= dr[1]) {
    // every value lies in [vlo, vhi], the ends of those digits' key
    // ranges; |v - med| computed as the selection computes it (v - med
    // rounds monotonically in v) is then at most max(vhi - med, med - vlo)
    constexpr int sh = int(sizeof(T)) * 8 - kSelBits;
    const T vlo = T(sel_val(uint64_t(dr[0]) << sh, (T*)nullptr));
    const T vhi = T(sel_val((uint64_t(dr[1]) << sh) | ((uint64_t(1) << sh) - 1), (T*)nullptr));
    const T cm = T(med);
    const double b = double(fmax(vhi - cm, cm - vlo));
    if (b > 0.0 && b <= 1.0e300 && double(kSelBins) / b < 1.0e300) bound = b;
  }
  return 1.4826 * block_median<T, NT>(v, off, stride, count, true, med, ws, nullptr, bound);
}

}  // namespace wh
