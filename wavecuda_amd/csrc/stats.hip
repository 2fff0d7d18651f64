// stats.hip -- thresholding and the smoothing statistics (gfx950).
//
//  * in-place hard/soft thresholding of DWT / MODWT details
//    (reference src/thresh.cpp:15-83, :142-157);
//  * exact median and MAD (reference src/utils.cpp:166-203) by MSB-first
//    radix selection over order-preserving integer keys of the values:
//    each grid pass builds an LDS-privatised histogram of one 11-bit digit
//    of the keys that still match the prefix found so far, and a one-block
//    scan picks the bucket holding the wanted rank; the few keys left after
//    the leading digits are compacted and one workgroup resolves the
//    remaining digits in LDS.  Both middle order
//    statistics are tracked at once; results are bit-exact (a median is an
//    order statistic; the final 0.5(a+b) and 1.4826 m are the reference's
//    operations);
//  * the universal threshold (reference src/thresh.cpp:170-201), whose
//    sqrt(2 ln n) factor is computed on the host by the C library so it
//    rounds like the reference;
//  * the CV interpolation error (reference src/thresh.cpp:344-384) as a
//    fixed-order two-stage reduction, so repeated runs are bit-identical.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "wh_common.h"
#include "wh_internal.h"
#include "wh_stats.h"

namespace wh {

// ---------------------------------------------------------------- threshold
template <typename T>
__global__ void thresh_dwt_kernel(T* __restrict__ y0, int64_t n, int J, Thr1 th) {
  T* y = y0 + int64_t(blockIdx.y) * n;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i == 0) continue;
    const int l = __builtin_ctzll(uint64_t(i));  // detail of 0-based level l
    if (l >= J || l < th.minlevel || l > th.maxlevel) continue;
    const T t = th.levels ? T(th.levels[l]) : T(th.value);
    y[i] = thresh_coef(y[i], t, th.hardness);
  }
}

template <typename T>
__global__ void thresh_modwt_kernel(T* __restrict__ xmod0, int64_t n, int J, Thr1 th) {
  T* xmod = xmod0 + int64_t(blockIdx.y) * 2 * n * J;
  const int64_t cnt = int64_t(th.maxlevel - th.minlevel + 1) * n;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < cnt; e += stride) {
    const int l = th.minlevel + int(e / n);
    const int64_t i = 2 * n * l + n + (e - int64_t(l - th.minlevel) * n);
    const T t = th.levels ? T(th.levels[l]) : T(th.value);
    xmod[i] = thresh_coef(xmod[i], t, th.hardness);
  }
}

int threshold_inplace(void* c, int64_t n, int64_t count, int J, int ttype, int dtype,
                      const wh_thresh* thr, hipStream_t st) {
  Thr1 th{thr->hardness, thr->minlevel, thr->maxlevel, thr->value, thr->levels};
  const int nt = 256;
  if (count < 1 || count > 65535) return set_error(WH_ERR_ARG, "threshold: signal count out of range");
  const int64_t work = (ttype == kDWT) ? n : int64_t(thr->maxlevel - thr->minlevel + 1) * n;
  const unsigned blocks = unsigned(std::min<int64_t>((work + nt - 1) / nt, 4096));
  const dim3 grid(blocks, unsigned(count));
  if (ttype == kDWT) {
    if (dtype == WH_F32)
      hipLaunchKernelGGL(thresh_dwt_kernel<float>, grid, dim3(nt), 0, st,
                         static_cast<float*>(c), n, J, th);
    else
      hipLaunchKernelGGL(thresh_dwt_kernel<double>, grid, dim3(nt), 0, st,
                         static_cast<double*>(c), n, J, th);
  } else {
    if (dtype == WH_F32)
      hipLaunchKernelGGL(thresh_modwt_kernel<float>, grid, dim3(nt), 0, st,
                         static_cast<float*>(c), n, J, th);
    else
      hipLaunchKernelGGL(thresh_modwt_kernel<double>, grid, dim3(nt), 0, st,
                         static_cast<double*>(c), n, J, th);
  }
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

// ------------------------------------------------------------ radix select

// order-preserving keys (larger value -> larger unsigned key)
__device__ __forceinline__ uint64_t okey(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_okey(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(int64_t(u));
}
__device__ __forceinline__ uint64_t okey(float v) {
  uint32_t u = __float_as_uint(v);
  return (u >> 31) ? uint64_t(~u) : uint64_t(u | 0x80000000u);
}

template <typename T>
__device__ __forceinline__ T value_from_key(uint64_t k);
template <>
__device__ __forceinline__ double value_from_key<double>(uint64_t k) {
  return from_okey(k);
}
template <>
__device__ __forceinline__ float value_from_key<float>(uint64_t k) {
  uint32_t kk = uint32_t(k);
  uint32_t u = (kk >> 31) ? (kk & 0x7fffffffu) : ~kk;
  return __uint_as_float(u);
}

// Element e of the selection source (see SelSrc).  Returns false when the
// position is not part of the selection (mode 2 filters by level).
template <typename T>
__device__ __forceinline__ bool sel_value(const SelSrc& s, int64_t e, T& v) {
  const T* __restrict__ b = static_cast<const T*>(s.base);
  if (s.mode == 0) {
    v = b[s.off + e];
    return true;
  }
  if (s.mode == 1) {  // MODWT detail blocks 0..nblocks-1
    const int64_t blk = e / s.n;
    v = b[2 * s.n * blk + s.n + (e - blk * s.n)];
    return true;
  }
  // mode 2: DWT details of 0-based levels [lmin, lmax] (position e, e >= 1)
  if (e == 0) return false;
  const int l = __builtin_ctzll(uint64_t(e));
  if (l >= s.J || l < s.lmin || l > s.lmax) return false;
  v = b[e];
  return true;
}

constexpr int kBits = 11;
constexpr int kBins = 1 << kBits;

__global__ void sel_init_kernel(SelState* st, int64_t count) {
  const int t = threadIdx.x;
  for (int i = t; i < 2 * kBins; i += blockDim.x) st->hist[i] = 0;
  if (t == 0) {
    st->prefix[0] = st->prefix[1] = 0;
    st->k[0] = (count - 1) / 2;
    st->k[1] = count / 2;
    st->ncand = 0;
    st->cmin[0] = st->cmin[1] = ~0ull;
    st->cmax[0] = st->cmax[1] = 0;
  }
}

// pass over all elements: histogram of digit [lo, hi) of the keys whose
// bits above hi equal the current prefix of rank r (r = 0, 1)
template <typename T>
__global__ __launch_bounds__(256) void sel_hist_kernel(SelSrc s, SelState* st, int hi, int lo,
                                                       int deviation) {
  __shared__ unsigned int h[2][kBins];
  for (int i = threadIdx.x; i < 2 * kBins; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  const uint64_t p0 = st->prefix[0], p1 = st->prefix[1];
  const bool same = (p0 == p1);
  const T med = deviation ? T(st->med) : T(0);
  const int64_t E = (s.mode == 2) ? s.n : s.count;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const uint64_t dmask = (uint64_t(1) << (hi - lo)) - 1;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += stride) {
    T v;
    if (!sel_value<T>(s, e, v)) continue;
    if (deviation) v = fabs(v - med);
    const uint64_t k = okey(v);
    const uint64_t top = (hi >= 64) ? 0 : (k >> hi);
    const unsigned d = unsigned((k >> lo) & dmask);
    if (top == p0) atomicAdd(&h[0][d], 1u);
    if (!same && top == p1) atomicAdd(&h[1][d], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) {
    if (h[0][i]) atomicAdd(&st->hist[i], (unsigned long long)h[0][i]);
    if (!same && h[1][i]) atomicAdd(&st->hist[kBins + i], (unsigned long long)h[1][i]);
  }
}

// one block of 256 threads: find the bucket of each rank, extend the
// prefixes, clear the histograms for the next pass
__global__ __launch_bounds__(256) void sel_scan_kernel(SelState* st, int hi, int lo) {
  __shared__ unsigned long long part[256];
  __shared__ int found_d;
  __shared__ long long found_below;
  const int t = threadIdx.x;
  constexpr int per = kBins / 256;
  const bool same = (st->prefix[0] == st->prefix[1]);
  for (int r = 0; r < 2; ++r) {
    const unsigned long long* hr = st->hist + ((same ? 0 : r) * kBins);
    unsigned long long loc = 0;
    for (int i = 0; i < per; ++i) loc += hr[t * per + i];
    part[t] = loc;
    __syncthreads();
    // inclusive scan of part (Hillis-Steele)
    for (int off = 1; off < 256; off <<= 1) {
      unsigned long long v = (t >= off) ? part[t - off] : 0;
      __syncthreads();
      part[t] += v;
      __syncthreads();
    }
    const long long k = st->k[r];
    const unsigned long long before = (t == 0) ? 0 : part[t - 1];
    if ((long long)before <= k && k < (long long)part[t]) {
      unsigned long long acc = before;
      for (int i = 0; i < per; ++i) {
        const unsigned long long c = hr[t * per + i];
        if ((long long)(acc + c) > k) {
          found_d = t * per + i;
          found_below = (long long)acc;
          break;
        }
        acc += c;
      }
    }
    __syncthreads();
    if (t == 0) {
      st->k[r] = k - found_below;
      st->prefix[r] = (st->prefix[r] << (hi - lo)) | uint64_t(found_d);
    }
    __syncthreads();
  }
  for (int i = t; i < 2 * kBins; i += 256) st->hist[i] = 0;
}

// After the grid passes: every key whose bits above `hi` equal either
// prefix is appended to the candidate buffer, one atomic per wave.  Past
// the buffer's capacity a wave stops appending (the count then exceeds
// the capacity and the finishing workgroup reads the data itself); the
// smallest and largest key under each prefix are kept either way, so a
// bucket of exact ties needs no further pass.
template <typename T>
__global__ __launch_bounds__(256) void sel_compact_kernel(SelSrc s, SelState* st, int hi,
                                                          int deviation) {
  __shared__ uint64_t red[4][4];
  const uint64_t p0 = st->prefix[0], p1 = st->prefix[1];
  const T med = deviation ? T(st->med) : T(0);
  const int64_t E = (s.mode == 2) ? s.n : s.count;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t mn0 = ~0ull, mx0 = 0, mn1 = ~0ull, mx1 = 0;
  bool full = false;
  for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < E; base += stride) {
    const int64_t e = base + threadIdx.x;
    bool c = false;
    uint64_t k = 0;
    T v;
    if (e < E && sel_value<T>(s, e, v)) {
      if (deviation) v = fabs(v - med);
      k = okey(v);
      const uint64_t top = k >> hi;
      if (top == p0) {
        mn0 = k < mn0 ? k : mn0;
        mx0 = k > mx0 ? k : mx0;
        c = true;
      }
      if (top == p1) {
        mn1 = k < mn1 ? k : mn1;
        mx1 = k > mx1 ? k : mx1;
        c = true;
      }
    }
    if (full) continue;
    const uint64_t mask = __ballot(c);
    if (!mask) continue;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const unsigned cnt = unsigned(__popcll(mask));
    unsigned long long old = 0;
    if (lane == leader) old = atomicAdd(&st->ncand, (unsigned long long)cnt);
    old = __shfl(old, leader, 64);
    if (c) {
      const unsigned long long idx = old + __popcll(mask & ((1ull << lane) - 1ull));
      if (idx < (unsigned long long)kSelCandCap) st->cand[idx] = k;
    }
    full = old + cnt > (unsigned long long)kSelCandCap;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a0 = __shfl_xor(mn0, o, 64), b0 = __shfl_xor(mx0, o, 64);
    const uint64_t a1 = __shfl_xor(mn1, o, 64), b1 = __shfl_xor(mx1, o, 64);
    mn0 = a0 < mn0 ? a0 : mn0;
    mx0 = b0 > mx0 ? b0 : mx0;
    mn1 = a1 < mn1 ? a1 : mn1;
    mx1 = b1 > mx1 ? b1 : mx1;
  }
  if (lane == 0) {
    red[wave][0] = mn0;
    red[wave][1] = mx0;
    red[wave][2] = mn1;
    red[wave][3] = mx1;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int j = threadIdx.x;
    uint64_t r = red[0][j];
    for (int w = 1; w < 4; ++w) r = (j & 1) ? (red[w][j] > r ? red[w][j] : r)
                                            : (red[w][j] < r ? red[w][j] : r);
    if (j & 1) {
      if (r != 0) atomicMax((unsigned long long*)&st->cmax[j >> 1], (unsigned long long)r);
    } else {
      if (r != ~0ull) atomicMin((unsigned long long*)&st->cmin[j >> 1], (unsigned long long)r);
    }
  }
}

constexpr int kFinNT = 1024;
constexpr size_t kFinLds = size_t(kSelCandCap) * 8 + size_t(2) * kBins * 4;

// inclusive prefix sum of v over the workgroup (kFinNT threads): within
// each wave by shuffles, then over the wave totals
__device__ __forceinline__ unsigned fin_scan(unsigned v, unsigned* wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  if (wave == 0) {
    unsigned w = lane < kFinNT / 64 ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < kFinNT / 64; o <<= 1) {
      const unsigned t = __shfl_up(w, o, 64);
      if (lane >= o) w += t;
    }
    if (lane < kFinNT / 64) wsum[lane] = w;
  }
  __syncthreads();
  return v + (wave ? wsum[wave - 1] : 0u);
}

// One workgroup resolves the digits below `hi0` and forms the median
// 0.5 (a + b) -- or the MAD 1.4826 median (times `factor` when > 0).  The
// keys under each middle rank's prefix come from the candidate buffer,
// held in LDS: up to kFinCount of them are ranked by direct counting;
// more take 11-bit digits, both ranks' bucket counts packed in one word
// (each at most kSelCandCap < 2^16) so that one scan serves both.  When
// the buffer overflowed, a bucket of one repeated value is that value, and
// otherwise the workgroup reads the values themselves.
constexpr int kFinCount = 256;

template <typename T>
__global__ __launch_bounds__(kFinNT) void sel_finish_kernel(SelSrc s, SelState* st, int hi0,
                                                            int deviation, double factor,
                                                            double* out) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* cand = reinterpret_cast<uint64_t*>(smem);
  unsigned* hist = reinterpret_cast<unsigned*>(cand + kSelCandCap);  // [2][kBins]
  __shared__ unsigned wsum[kFinNT / 64];
  __shared__ uint64_t pre[2];
  __shared__ long long ks[2];
  __shared__ int dsel[2];
  __shared__ long long below[2];
  const int tid = threadIdx.x;
  const unsigned long long nall = hi0 > 0 ? st->ncand : 0;
  const bool in_lds = nall <= (unsigned long long)kSelCandCap;
  const int nc = in_lds ? int(nall) : 0;
  // overflow with each bucket holding a single key value: those are the keys
  const bool ties = !in_lds && st->cmin[0] == st->cmax[0] && st->cmin[1] == st->cmax[1];
  const bool count = in_lds && hi0 > 0 && nc <= kFinCount;
  if (tid == 0) {
    pre[0] = ties ? st->cmin[0] : st->prefix[0];
    pre[1] = ties ? st->cmin[1] : st->prefix[1];
    ks[0] = st->k[0];
    ks[1] = st->k[1];
  }
  for (int i = tid; i < nc; i += kFinNT) cand[i] = st->cand[i];
  const T med = deviation ? T(st->med) : T(0);
  const int64_t E = (s.mode == 2) ? s.n : s.count;
  __syncthreads();
  if (count) {
    // rank of key i under its prefix = keys of the same prefix below it;
    // threads holding equal keys write the same value
    uint64_t p0 = pre[0], p1 = pre[1];
    long long k0 = ks[0], k1 = ks[1];
    __syncthreads();
    if (tid < nc) {
      const uint64_t ki = cand[tid], top = ki >> hi0;
      unsigned less = 0, eq = 0;
      for (int j = 0; j < nc; ++j) {
        const uint64_t kj = cand[j];
        const bool same = (kj >> hi0) == top;
        less += (same && kj < ki) ? 1u : 0u;
        eq += (same && kj == ki) ? 1u : 0u;
      }
      if (top == p0 && less <= k0 && k0 < (long long)(less + eq)) pre[0] = ki;
      if (top == p1 && less <= k1 && k1 < (long long)(less + eq)) pre[1] = ki;
    }
    __syncthreads();
  }
  for (int hi = (ties || count) ? 0 : hi0; hi > 0;) {
    const int lo = hi > kBits ? hi - kBits : 0;
    for (int i = tid; i < 2 * kBins; i += kFinNT) hist[i] = 0;
    __syncthreads();
    const uint64_t p0 = pre[0], p1 = pre[1];
    const bool same = (p0 == p1);
    const uint64_t dmask = (uint64_t(1) << (hi - lo)) - 1;
    constexpr int per = kBins / kFinNT;
    if (in_lds) {
      for (int i = tid; i < nc; i += kFinNT) {
        const uint64_t k = cand[i], top = k >> hi;
        const unsigned inc = (top == p0 ? 1u : 0u) + (top == p1 ? 0x10000u : 0u);
        if (inc) atomicAdd(&hist[unsigned((k >> lo) & dmask)], inc);
      }
      __syncthreads();
      unsigned loc = 0;
#pragma unroll
      for (int i = 0; i < per; ++i) loc += hist[tid * per + i];
      const unsigned incl = fin_scan(loc, wsum);
      const unsigned before = incl - loc;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int sh = 16 * r;
        const long long k = ks[r];
        const long long b = (before >> sh) & 0xffffu, e = (incl >> sh) & 0xffffu;
        if (b <= k && k < e) {
          long long acc = b;
          for (int i = 0; i < per; ++i) {
            const long long c = (hist[tid * per + i] >> sh) & 0xffffu;
            if (acc + c > k) {
              dsel[r] = tid * per + i;
              below[r] = acc;
              break;
            }
            acc += c;
          }
        }
      }
      __syncthreads();
      if (tid < 2) {
        ks[tid] -= below[tid];
        pre[tid] = (pre[tid] << (hi - lo)) | uint64_t(dsel[tid]);
      }
      __syncthreads();
    } else {
      for (int64_t e = tid; e < E; e += kFinNT) {
        T v;
        if (!sel_value<T>(s, e, v)) continue;
        if (deviation) v = fabs(v - med);
        const uint64_t k = okey(v), top = k >> hi;
        const unsigned d = unsigned((k >> lo) & dmask);
        if (top == p0) atomicAdd(&hist[d], 1u);
        if (!same && top == p1) atomicAdd(&hist[kBins + d], 1u);
      }
      __syncthreads();
      for (int r = 0; r < 2; ++r) {
        const unsigned* hr = hist + (same ? 0 : r) * kBins;
        const long long k = ks[r];
        unsigned loc = 0;
#pragma unroll
        for (int i = 0; i < per; ++i) loc += hr[tid * per + i];
        const unsigned incl = fin_scan(loc, wsum);
        const long long before = (long long)(incl - loc);
        if (before <= k && k < (long long)incl) {
          long long acc = before;
          for (int i = 0; i < per; ++i) {
            const unsigned c = hr[tid * per + i];
            if (acc + c > k) {
              dsel[r] = tid * per + i;
              below[r] = acc;
              break;
            }
            acc += c;
          }
        }
        __syncthreads();
        if (tid == 0) {
          ks[r] = k - below[r];
          pre[r] = (pre[r] << (hi - lo)) | uint64_t(dsel[r]);
        }
        __syncthreads();
      }
    }
    hi = lo;
  }
  if (tid == 0) {
    const double a = double(value_from_key<T>(pre[0]));
    const double b = double(value_from_key<T>(pre[1]));
    const double m = 0.5 * (a + b);
    if (!deviation) {
      st->med = m;
      if (out) *out = m;
    } else {
      st->mad = 1.4826 * m;
      if (out) *out = (factor > 0.0) ? st->mad * factor : st->mad;
    }
  }
}

// Number of 11-bit digits resolved by grid passes before the keys still
// in play are compacted.  After two digits the bucket of a middle rank
// holds the values within about 2^-13 (FP32 keys) or 2^-10 (FP64 keys) of
// it; beyond 2^22 values (either type) a third grid digit narrows it, so
// the keys left (about 2^22 * 2^-13 of a Gaussian-like sample for FP32)
// still fit the finishing workgroup's LDS instead of overflowing into its
// slow path that reads every value.  Beyond 2^31 values every digit is a
// grid pass (the finishing workgroup counts in 32 bits).
inline int grid_digits(int key_bits, int64_t E) {
  if (E >= (int64_t(1) << 31)) return (key_bits + kBits - 1) / kBits;
  return E <= (int64_t(1) << 22) ? 2 : 3;
}

// Grid histogram passes over the leading digits, then a compaction of the
// keys still in play and one workgroup that resolves the remaining digits
// in LDS: seven launches per median (nine for FP64 beyond 2^22 values)
// instead of thirteen (FP64) or eight (FP32), three passes over the values
// (four) instead of six (three).
template <typename T>
int select_median_t(const SelSrc& s, bool deviation, SelState* st, double factor, double* out,
                    hipStream_t stream) {
  const int key_bits = sizeof(T) * 8;
  allow_dynamic_lds(reinterpret_cast<const void*>(sel_finish_kernel<T>));
  hipLaunchKernelGGL(sel_init_kernel, dim3(1), dim3(256), 0, stream, st, s.count);
  const int64_t E = (s.mode == 2) ? s.n : s.count;
  const int cus = std::max(1, device_info().num_cus);
  const unsigned blocks = unsigned(std::max<int64_t>(1, std::min<int64_t>((E + 2047) / 2048, 4 * cus)));
  const int gd = grid_digits(key_bits, E);
  int hi = key_bits;
  for (int d = 0; d < gd; ++d) {
    const int lo = hi > kBits ? hi - kBits : 0;
    hipLaunchKernelGGL(sel_hist_kernel<T>, dim3(blocks), dim3(256), 0, stream, s, st, hi, lo,
                       deviation ? 1 : 0);
    hipLaunchKernelGGL(sel_scan_kernel, dim3(1), dim3(256), 0, stream, st, hi, lo);
    hi = lo;
  }
  if (hi > 0)
    hipLaunchKernelGGL(sel_compact_kernel<T>, dim3(blocks), dim3(256), 0, stream, s, st, hi,
                       deviation ? 1 : 0);
  hipLaunchKernelGGL(sel_finish_kernel<T>, dim3(1), dim3(kFinNT), hi > 0 ? kFinLds : 0, stream,
                     s, st, hi, deviation ? 1 : 0, factor, out);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

int select_median(const SelSrc& s, int dtype, bool deviation, SelState* st, double factor,
                  double* out, hipStream_t stream) {
  if (s.count <= 0) return set_error(WH_ERR_ARG, "median of an empty selection");
  return dtype == WH_F32 ? select_median_t<float>(s, deviation, st, factor, out, stream)
                         : select_median_t<double>(s, deviation, st, factor, out, stream);
}

// MAD of the selection; result (times `factor` when > 0) to device `out`
int select_mad(const SelSrc& s, int dtype, SelState* st, double factor, double* out,
               hipStream_t stream) {
  int rc = select_median(s, dtype, false, st, 0.0, nullptr, stream);
  if (rc) return rc;
  return select_median(s, dtype, true, st, factor, out, stream);
}

// ------------------------------------------------- cross-validation search
// The scalar steps of the reference's golden-section loop
// (src/thresh.cpp:296-330), with products and sums rounded separately as
// in its host build, so the visited thresholds are the same doubles.  One
// thread: take the error m of the pending point, advance the bracket and
// write the next point into levels[0..J).
__device__ __noinline__ void cv_advance(CvState* cs, double m, double* levels, int J,
                                        double tol) {
#pragma clang fp contract(off)
  if (cs->stop) return;
  const double R = 0.61803399, C = 1. - R;
  if (cs->pending == 1)
    cs->m1 = m;
  else
    cs->m2 = m;
  const int ne = ++cs->nevals;
  double next;
  if (ne == 1) {  // the first two evaluations: t1, then t2
    cs->pending = 2;
    next = cs->t2;
  } else {
    if (ne > 2) cs->iter++;  // the loop body that ordered this evaluation ends
    if (!(fabs(cs->t3 - cs->t0) > tol * (cs->t1 + cs->t2)) || cs->iter > 50) {
      cs->stop = 1;
      return;
    }
    if (cs->m2 < cs->m1) {
      cs->t0 = cs->t1;
      cs->t1 = cs->t2;
      cs->t2 = R * cs->t2 + C * cs->t3;
      cs->m1 = cs->m2;
      cs->pending = 2;
      next = cs->t2;
    } else {
      cs->t3 = cs->t2;
      cs->t2 = cs->t1;
      cs->t1 = R * cs->t1 + C * cs->t0;
      cs->m2 = cs->m1;
      cs->pending = 1;
      next = cs->t1;
    }
  }
  for (int l = 0; l < J; ++l) levels[l] = next;
}

__global__ void cv_init_k(CvState* cs, const double* tc, double* levels, int J) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double C = 1. - 0.61803399;
  const double t3 = *tc, tb = t3 / 2.;
  cs->t0 = 0.;
  cs->t2 = tb;
  cs->t1 = tb - C * (tb - 0.);
  cs->t3 = t3;
  cs->m1 = cs->m2 = 0.;
  cs->iter = 0;
  cs->pending = 1;
  cs->nevals = 0;
  cs->stop = 0;
  for (int l = 0; l < J; ++l) levels[l] = cs->t1;
}

int cv_init(CvState* cs, const double* tc, double* levels, int J, hipStream_t st) {
  hipLaunchKernelGGL(cv_init_k, dim3(1), dim3(64), 0, st, cs, tc, levels, J);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

// --------------------------------------------------------- interp_mse
template <typename T>
__global__ __launch_bounds__(256) void interp_mse_partial(const T* __restrict__ xn,
                                                          const T* __restrict__ ye,
                                                          const T* __restrict__ yo, int64_t len,
                                                          int64_t per_block, double* partial) {
  __shared__ double red[256];
  const int64_t lenh = len / 2;
  const int64_t j0 = int64_t(blockIdx.x) * per_block;
  const int64_t j1 = std::min(lenh, j0 + per_block);
  double acc = 0.0;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const int64_t i = 2 * j;
    double yie, yio;
    if (i == 0) {
      yie = double(yo[0]);
      yio = 0.5 * (double(ye[0]) + double(ye[1]));
    } else if (i == len - 2) {
      yio = double(ye[lenh - 1]);
      yie = 0.5 * (double(yo[j]) + double(yo[j - 1]));
    } else {
      yio = 0.5 * (double(ye[j]) + double(ye[j + 1]));
      yie = 0.5 * (double(yo[j]) + double(yo[j - 1]));
    }
    const double a = yie - double(xn[i]), b = yio - double(xn[i + 1]);
    acc += a * a + b * b;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_partials(const double* partial, int np, double scale,
                                                    double* out, CvArgs cv) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = red[0] * scale;
    *out = v;
    if (cv.cs) cv_advance(cv.cs, v, cv.levels, cv.J, cv.tol);
  }
}

int interp_mse(const void* xn, const void* ye, const void* yo, int64_t len, int dtype,
               double* partial_buf, double* out, hipStream_t st, const CvArgs& cv) {
  const int64_t lenh = len / 2;
  const int64_t per_block = 512;
  const int np = int(std::min<int64_t>((lenh + per_block - 1) / per_block, kMaxPartials));
  const int64_t pb = (lenh + np - 1) / np;
  if (dtype == WH_F32)
    hipLaunchKernelGGL(interp_mse_partial<float>, dim3(np), dim3(256), 0, st,
                       static_cast<const float*>(xn), static_cast<const float*>(ye),
                       static_cast<const float*>(yo), len, pb, partial_buf);
  else
    hipLaunchKernelGGL(interp_mse_partial<double>, dim3(np), dim3(256), 0, st,
                       static_cast<const double*>(xn), static_cast<const double*>(ye),
                       static_cast<const double*>(yo), len, pb, partial_buf);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, partial_buf, np, 0.5, out, cv);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

// ------------------------------------------------------------- helpers
template <typename T>
__global__ void split_even_odd(const T* __restrict__ x, T* __restrict__ e, T* __restrict__ o,
                               int64_t lenh) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < lenh; j += stride) {
    e[j] = x[2 * j];
    o[j] = x[2 * j + 1];
  }
}

int split_halves(const void* x, void* e, void* o, int64_t len, int dtype, hipStream_t st) {
  const int64_t lenh = len / 2;
  const unsigned blocks = unsigned(std::max<int64_t>(1, std::min<int64_t>((lenh + 255) / 256, 4096)));
  if (dtype == WH_F32)
    hipLaunchKernelGGL(split_even_odd<float>, dim3(blocks), dim3(256), 0, st,
                       static_cast<const float*>(x), static_cast<float*>(e),
                       static_cast<float*>(o), lenh);
  else
    hipLaunchKernelGGL(split_even_odd<double>, dim3(blocks), dim3(256), 0, st,
                       static_cast<const double*>(x), static_cast<double*>(e),
                       static_cast<double*>(o), lenh);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

__global__ void fill_levels(double* dst, const double* src, double scale, int J) {
  const int t = threadIdx.x;
  if (t < J) dst[t] = (*src) * scale;
}

int broadcast_threshold(double* dst, const double* src, double scale, int J, hipStream_t st) {
  hipLaunchKernelGGL(fill_levels, dim3(1), dim3(64), 0, st, dst, src, scale, J);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

}  // namespace wh
