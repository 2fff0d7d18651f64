// batch_kern.h -- device side of the batched smoothing pipeline.
//
// Thousands of independent signals of power-of-two lengths are smoothed
// (forward transform -> hard/soft universal threshold -> inverse) by
// persistent workgroups that take whole signals from a device work queue:
// an atomic counter over an item list sorted longest-first, so the large
// signals start first and the small ones fill the tail.  One launch serves
// every transform, filter and length of the batch: each item dispatches to
// a body specialised for its filter (and, for the MODWT, its depth, which is
// the fused group depth), so the chunk geometry and tap arithmetic are
// compile-time.  Every workgroup owns a slice of scratch for the
// coefficients of the signal it is working on; nothing of the pipeline
// returns to the host.
//
//   MODWT: the fused sliding-window forward (modwt_fwd_body) writes the J
//   levels into scratch, block_mad takes the exact MAD of D_1 (the
//   reference's univ_thresh(w, 0, 0): src/thresh.cpp:170-201,
//   src/rfunctions.cu:82-87), and the fused inverse (modwt_inv_lin_body)
//   thresholds every detail level as it reads it.
//   DWT: levels run on compact sequences, through scratch while they are
//   longer than the LDS tile and in LDS afterwards; the MAD is taken over
//   the level-1 details (all of x[1..n-1] when J = 1, as in the reference)
//   and the inverse thresholds the details as it reads them.
#pragma once

#include "blocksel.h"
#include "dwt_dev.h"
#include "modwt_kern.h"

namespace wh {

// transform and filter of an item in one word
__host__ __device__ constexpr int batch_kind(int ttype, int filt) { return (ttype << 3) | filt; }
__host__ __device__ constexpr int kind_ttype(int kind) { return kind >> 3; }
__host__ __device__ constexpr int kind_filter(int kind) { return kind & 7; }

// One signal of the batch.
struct BatchItem {
  int64_t off;    // first sample in the packed signal buffers
  double factor;  // sqrt(2 ln n) (sqrt(2 ln(n-1)) for a one-level DWT)
  int32_t n;      // length (power of two)
  int32_t J;      // levels
  int32_t kind;   // batch_kind(ttype, filter)
  int32_t sig;    // signal index (threshold output slot)
};

struct BatchArgs {
  const BatchItem* items;
  int count;
  int* counter;  // [0] work-queue head, [1] finished workgroups
  const void* x;
  void* y;
  double* thresh_out;  // optional, one per signal
  void* scratch;
  int64_t scratch_stride;  // elements per workgroup
  int hardness;
};

__device__ __forceinline__ int batch_next(int* counter, int* sh) {
  __syncthreads();
  if (threadIdx.x == 0) *sh = atomicAdd(counter, 1);
  __syncthreads();
  return *sh;
}

// phase boundary inside a workgroup when the phases exchange data through
// global memory (the workgroup's own scratch slice).  Only this
// workgroup's waves, all on one CU, read what they wrote, so the barrier's
// workgroup-scope release/acquire is enough; a device-scope fence would
// write back and invalidate the L2 of the workgroup's XCD at every item.
__device__ __forceinline__ void batch_phase() { __syncthreads(); }

// MODWT item, forward: the fused forward of all J = G levels into scratch
// (D_j and S_J).  Chunk width M = min(2 NT, n); the warm-up covers the
// coarsest level's reach.
template <typename T, int FILT, int NT, int G>
__device__ __forceinline__ ModwtArgs batch_modwt_args(const BatchArgs& b, const BatchItem& I,
                                                      T* scr) {
  constexpr int L = Filter<FILT>::L;
  const int64_t n = I.n;
  ModwtArgs a{};
  a.n = n;
  a.P = n;
  a.j0 = 0;
  a.g = G;
  a.M = int(n < 2 * NT ? n : 2 * NT);
  const int64_t reach = int64_t(L - 1) * ((int64_t(1) << G) - 1);
  a.warm = int((reach + a.M - 1) / a.M * a.M);
  a.nchunks = int((n + a.warm) / a.M);
  a.in = static_cast<const T*>(b.x) + I.off;
  a.xmod = scr;
  return a;
}

template <typename T, int FILT, int NT, int G>
__device__ __forceinline__ void batch_modwt_fwd(const BatchArgs& b, const BatchItem& I, T* scr,
                                                unsigned char* smem, unsigned* pre) {
  const ModwtArgs a = batch_modwt_args<T, FILT, NT, G>(b, I, scr);
  modwt_fwd_body<T, FILT, 1, NT, G, 0, kStoreSmoothHist>(a, 0, 0, smem, pre);
}

// MODWT item, inverse: the fused inverse into y, thresholding every detail
// level at t as it reads it
template <typename T, int FILT, int NT, int G>
__device__ __forceinline__ void batch_modwt_inv(const BatchArgs& b, const BatchItem& I, T* scr,
                                                unsigned char* smem, double t) {
  ModwtArgs a = batch_modwt_args<T, FILT, NT, G>(b, I, scr);
  const int64_t n = I.n;
  a.stop = scr + 2 * n * (G - 1);
  a.out = static_cast<T*>(b.y) + I.off;
  a.th_on = 1;
  a.hardness = b.hardness;
  a.minlevel = 0;
  a.maxlevel = G - 1;
  a.th_value = t;
  a.th_levels = nullptr;
  modwt_inv_lin_body<T, FILT, NT, G>(a, 0, smem);
}

// LDS elements per ping-pong buffer of the DWT tail (64 KB in FP64, 32 KB
// in FP32, so that several FP32 workgroups share a CU)
template <typename T>
constexpr int batch_lds_elems() {
  return 4096;
}

// FP64 512-thread workgroups keep a first-digit histogram of the median's
// values (2048 counts) at this LDS offset, beyond the forward's rings (at
// most 80 KB at J = 8), the DWT's LDS levels (64 KB) and the selection's
// workspace; the plan sizes the launch's LDS to cover it.  Other shapes
// would need more LDS per workgroup than they have to spare.
constexpr size_t kBatchPreOff = size_t(96) << 10;
constexpr size_t kBatchPreEnd = kBatchPreOff + size_t(kSelBins) * 4;
// LDS the histogram needs for NT threads and esz-byte elements (0: none)
__host__ __device__ constexpr size_t batch_pre_lds(int NT, size_t esz) {
  return (esz == 8 && NT == 512) ? kBatchPreEnd : 0;
}
template <typename T, int NT>
__host__ __device__ constexpr bool batch_pre() {
  return batch_pre_lds(NT, sizeof(T)) != 0;
}

// DWT item coefficients in scratch C, level by level and each level
// contiguous: d_j at C[n - n/2^{j-1} ..) (n/2^j values), then s_J; for
// J = 1, s_1[0] goes last, so that C[0 .. n-1) holds d_1 and s_1[1..]: the
// values of the reference's interleaved layout at positions 1..n-1, which
// its one-level universal threshold reads (src/thresh.cpp:170-201).
__device__ __forceinline__ int64_t batch_dwt_doff(int64_t n, int j) { return n - (n >> (j - 1)); }
__device__ __forceinline__ int64_t batch_dwt_sidx(int64_t n, int J, int64_t m) {
  if (J == 1) return m == 0 ? n - 1 : n / 2 + m - 1;
  return n - (n >> J) + m;
}

// DWT item, forward: levels run on compact sequences, through scratch while
// longer than the LDS tile and in LDS afterwards; the coefficients end in
// scratch C as above (n values, gA/gB n/2 each behind it).
template <typename T, int FILT, int NT>
__device__ __forceinline__ void batch_dwt_fwd(const BatchArgs& b, const BatchItem& I, T* C,
                                              unsigned char* smem, unsigned* pre) {
  constexpr int L = Filter<FILT>::L;
  constexpr int kBatchLds = batch_lds_elems<T>();
  constexpr int kUp = 2;  // global-level output pairs of pairs per thread in flight
  typedef T T2 __attribute__((ext_vector_type(2)));
  T* lA = reinterpret_cast<T*>(smem);
  T* lB = lA + kBatchLds;
  const int tid = threadIdx.x;
  {
    const int64_t n = I.n;
    const int J = I.J;
    T* gA = C + n;          // n/2
    T* gB = gA + n / 2;     // n/2
    const T* x = static_cast<const T*>(b.x) + I.off;
    T* y = static_cast<T*>(b.y) + I.off;

    // ---- forward: global levels while the input is longer than the tile
    const T* src = x;
    int64_t N = n;
    int j = 1;
    bool toA = true;
    while (j <= J && N > kBatchLds) {
      T* dst = toA ? gA : gB;
      T* dj = C + batch_dwt_doff(n, j);
      // each thread takes outputs m = 2p and 2p + 1, whose taps 4p ..
      // 4p + L + 1 come as L / 2 + 1 aligned element pairs (N is a multiple
      // of 4, so no pair straddles the wrap), and stores them as 16-byte
      // pairs; kUp such pairs per thread in flight, clamped p for
      // unconditional loads
      const int64_t np = N / 4;
      for (int64_t p0 = tid; p0 < np; p0 += int64_t(NT) * kUp) {
        T v[kUp][L + 2];
#pragma unroll
        for (int u = 0; u < kUp; ++u) {
          const int64_t p = p0 + int64_t(u) * NT;
          const int64_t pc = (p < np) ? p : p0;
#pragma unroll
          for (int k = 0; k < L + 2; k += 2) {
            const T2 pr = *reinterpret_cast<const T2*>(src + ((4 * pc + k) & (N - 1)));
            v[u][k] = pr.x;
            v[u][k + 1] = pr.y;
          }
        }
#pragma unroll
        for (int u = 0; u < kUp; ++u) {
          const int64_t p = p0 + int64_t(u) * NT;
          if (p < np) {
            T s0, d0, s1, d1;
            fwd_pair<T, FILT>(v[u], s0, d0);
            fwd_pair<T, FILT>(v[u] + 2, s1, d1);
            *reinterpret_cast<T2*>(dst + 2 * p) = T2{s0, s1};
            *reinterpret_cast<T2*>(dj + 2 * p) = T2{d0, d1};
            if (pre && j == 1) {
              atomicAdd(&pre[sel_digit0(d0)], 1u);
              atomicAdd(&pre[sel_digit0(d1)], 1u);
            }
          }
        }
      }
      batch_phase();
      src = dst;
      toA = !toA;
      N >>= 1;
      ++j;
    }
    // remaining levels in LDS (or, when every level ran above, s_J itself)
    T* cur = lA;
    T* nxt = lB;
    const bool all_global = (j > J);
    if (all_global) {
      for (int64_t m = tid; m < N; m += NT) C[batch_dwt_sidx(n, J, m)] = src[m];
    } else {
      for (int64_t i = tid; i < N; i += NT) lA[i] = src[i];
    }
    __syncthreads();
    for (; j <= J; ++j) {
      const int P = int(N / 2);
      for (int m = tid; m < P; m += NT) {
        T v[L];
#pragma unroll
        for (int k = 0; k < L; k += 2) {
          const T2 pr = *reinterpret_cast<const T2*>(cur + ((2 * m + k) & (N - 1)));
          v[k] = pr.x;
          v[k + 1] = pr.y;
        }
        T s, d;
        fwd_pair<T, FILT>(v, s, d);
        nxt[m] = s;
        C[batch_dwt_doff(n, j) + m] = d;
        if (pre && j == 1) atomicAdd(&pre[sel_digit0(d)], 1u);
      }
      __syncthreads();
      T* t = cur;
      cur = nxt;
      nxt = t;
      N = P;
    }
    if (!all_global)
      for (int64_t m = tid; m < N; m += NT) C[batch_dwt_sidx(n, J, m)] = cur[m];
    batch_phase();
  }
}

// DWT item, inverse: from the interleaved coefficients in scratch C, the
// thresholding inverse into y (levels J.. in LDS while the output fits the
// tile, the long levels through gA/gB)
template <typename T, int FILT, int NT>
__device__ __forceinline__ void batch_dwt_inv(const BatchArgs& b, const BatchItem& I, T* C,
                                              unsigned char* smem, double t) {
  constexpr int L = Filter<FILT>::L;
  constexpr int kBatchLds = batch_lds_elems<T>();
  constexpr int kUp = 2;  // global-level output pairs of pairs per thread in flight
  typedef T T2 __attribute__((ext_vector_type(2)));
  T* lA = reinterpret_cast<T*>(smem);
  T* lB = lA + kBatchLds;
  const int tid = threadIdx.x;
  {
    const int64_t n = I.n;
    const int J = I.J;
    T* gA = C + n;          // n/2
    T* gB = gA + n / 2;     // n/2
    T* y = static_cast<T*>(b.y) + I.off;
    const T tt = T(t);
    int64_t N;
    T* cur;
    T* nxt;
    int j;

    // ---- inverse: levels J.. in LDS while the output fits the tile
    N = n >> J;  // length of s_J
    const bool lds_tail = (2 * N <= kBatchLds);
    if (lds_tail) {
      for (int64_t m = tid; m < N; m += NT) lA[m] = C[batch_dwt_sidx(n, J, m)];
    } else {
      for (int64_t m = tid; m < N; m += NT) gA[m] = C[batch_dwt_sidx(n, J, m)];
    }
    batch_phase();
    cur = lA;
    nxt = lB;
    j = J;
    while (j >= 1 && 2 * N <= kBatchLds) {
      for (int m = tid; m < int(N); m += NT) {
        T sv[L / 2], dv[L / 2];
#pragma unroll
        for (int q = 0; q < L / 2; ++q) {
          const int mm = (m - q) & int(N - 1);
          sv[q] = cur[mm];
          dv[q] = thresh_coef(C[batch_dwt_doff(n, j) + mm], tt, b.hardness);
        }
        T e, o;
        inv_pair<T, FILT>(sv, dv, e, o);
        *reinterpret_cast<T2*>(nxt + 2 * m) = T2{e, o};
      }
      __syncthreads();
      T* tp = cur;
      cur = nxt;
      nxt = tp;
      N *= 2;
      --j;
    }
    if (j == 0) {
      for (int64_t i = tid; i < n; i += NT) y[i] = cur[i];
    } else {
      // hand s_j to global (already there when no level ran in LDS) and
      // finish the long levels there
      if (lds_tail) {
        for (int64_t i = tid; i < N; i += NT) gA[i] = cur[i];
        batch_phase();
      }
      const T* s = gA;
      bool inA = true;
      // each thread takes the outputs of m = 2p and 2p + 1: their taps
      // 2p - H + 1 .. 2p + 1 of s_j and d_j come as K aligned element pairs
      // (N is even here, so no pair straddles the wrap), and the four
      // outputs go out as two 16-byte pair stores
      constexpr int H = L / 2, K = (H + 2) / 2;
      for (; j >= 1; --j) {
        T* out = (j == 1) ? y : (inA ? gB : gA);
        const T* dj = C + batch_dwt_doff(n, j);
        const int64_t np = N / 2;
        for (int64_t p0 = tid; p0 < np; p0 += int64_t(NT) * kUp) {
          T ws[kUp][2 * K], wd[kUp][2 * K];
#pragma unroll
          for (int u = 0; u < kUp; ++u) {
            const int64_t p = p0 + int64_t(u) * NT;
            const int64_t pc = (p < np) ? p : p0;  // clamped: unconditional loads
#pragma unroll
            for (int k = 0; k < K; ++k) {
              const int64_t i = (2 * pc - 2 * K + 2 + 2 * k) & (N - 1);
              const T2 a = *reinterpret_cast<const T2*>(s + i);
              const T2 d = *reinterpret_cast<const T2*>(dj + i);
              ws[u][2 * k] = a.x;
              ws[u][2 * k + 1] = a.y;
              wd[u][2 * k] = d.x;
              wd[u][2 * k + 1] = d.y;
            }
          }
#pragma unroll
          for (int u = 0; u < kUp; ++u) {
            const int64_t p = p0 + int64_t(u) * NT;
            if (p < np) {
#pragma unroll
              for (int i = 0; i < 2 * K; ++i) wd[u][i] = thresh_coef(wd[u][i], tt, b.hardness);
              // output m = 2p + r reads s_j[m - q] = w[2K - 2 + r - q]
              T sv[H], dv[H], e0, o0, e1, o1;
#pragma unroll
              for (int q = 0; q < H; ++q) {
                sv[q] = ws[u][2 * K - 2 - q];
                dv[q] = wd[u][2 * K - 2 - q];
              }
              inv_pair<T, FILT>(sv, dv, e0, o0);
#pragma unroll
              for (int q = 0; q < H; ++q) {
                sv[q] = ws[u][2 * K - 1 - q];
                dv[q] = wd[u][2 * K - 1 - q];
              }
              inv_pair<T, FILT>(sv, dv, e1, o1);
              *reinterpret_cast<T2*>(out + 4 * p) = T2{e0, o0};
              *reinterpret_cast<T2*>(out + 4 * p + 2) = T2{e1, o1};
            }
          }
        }
        batch_phase();
        s = out;
        inA = !inA;
        N *= 2;
      }
    }
  }
  batch_phase();
}

template <typename T, int FILT, int NT>
__device__ __forceinline__ void batch_fwd(const BatchArgs& b, const BatchItem& I, T* scr,
                                          unsigned char* smem, unsigned* pre) {
  if (kind_ttype(I.kind) == kDWT) {
    batch_dwt_fwd<T, FILT, NT>(b, I, scr, smem, pre);
    return;
  }
  switch (I.J) {
    case 1: batch_modwt_fwd<T, FILT, NT, 1>(b, I, scr, smem, pre); break;
    case 2: batch_modwt_fwd<T, FILT, NT, 2>(b, I, scr, smem, pre); break;
    case 3: batch_modwt_fwd<T, FILT, NT, 3>(b, I, scr, smem, pre); break;
    case 4: batch_modwt_fwd<T, FILT, NT, 4>(b, I, scr, smem, pre); break;
    case 5: batch_modwt_fwd<T, FILT, NT, 5>(b, I, scr, smem, pre); break;
    case 6: batch_modwt_fwd<T, FILT, NT, 6>(b, I, scr, smem, pre); break;
    case 7: batch_modwt_fwd<T, FILT, NT, 7>(b, I, scr, smem, pre); break;
    default: batch_modwt_fwd<T, FILT, NT, 8>(b, I, scr, smem, pre); break;
  }
}

template <typename T, int FILT, int NT>
__device__ __forceinline__ void batch_inv(const BatchArgs& b, const BatchItem& I, T* scr,
                                          unsigned char* smem, double t) {
  if (kind_ttype(I.kind) == kDWT) {
    batch_dwt_inv<T, FILT, NT>(b, I, scr, smem, t);
    return;
  }
  switch (I.J) {
    case 1: batch_modwt_inv<T, FILT, NT, 1>(b, I, scr, smem, t); break;
    case 2: batch_modwt_inv<T, FILT, NT, 2>(b, I, scr, smem, t); break;
    case 3: batch_modwt_inv<T, FILT, NT, 3>(b, I, scr, smem, t); break;
    case 4: batch_modwt_inv<T, FILT, NT, 4>(b, I, scr, smem, t); break;
    case 5: batch_modwt_inv<T, FILT, NT, 5>(b, I, scr, smem, t); break;
    case 6: batch_modwt_inv<T, FILT, NT, 6>(b, I, scr, smem, t); break;
    case 7: batch_modwt_inv<T, FILT, NT, 7>(b, I, scr, smem, t); break;
    default: batch_modwt_inv<T, FILT, NT, 8>(b, I, scr, smem, t); break;
  }
}

// The persistent pipeline: every workgroup takes items until the queue is
// empty; the last workgroup to leave rewinds the queue, so the launch can be
// replayed (a captured graph) without a memset node.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void batch_smooth_k(BatchArgs b) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int item_sh;
  T* scr = static_cast<T*>(b.scratch) + int64_t(blockIdx.x) * b.scratch_stride;
  for (;;) {
    const int it = batch_next(b.counter, &item_sh);
    if (it >= b.count) break;
    const BatchItem I = b.items[it];
    const int64_t n = I.n;
    const bool dwt = kind_ttype(I.kind) == kDWT;
    // the forward counts the first radix digit of the values the median
    // selects from (the level-1 details; not for a one-level DWT, whose
    // selection also takes s_1[1..]) into LDS the forward and the
    // selection leave free
    unsigned* pre = nullptr;
    if constexpr (batch_pre<T, NT>()) {
      if (!(dwt && I.J == 1)) {
        pre = reinterpret_cast<unsigned*>(smem + kBatchPreOff);
        for (int i = threadIdx.x; i < kSelBins; i += NT) pre[i] = 0;
        __syncthreads();
      }
    }
    switch (kind_filter(I.kind)) {
      case kHAAR: batch_fwd<T, kHAAR, NT>(b, I, scr, smem, pre); break;
      case kD4: batch_fwd<T, kD4, NT>(b, I, scr, smem, pre); break;
      case kC6: batch_fwd<T, kC6, NT>(b, I, scr, smem, pre); break;
      default: batch_fwd<T, kLA8, NT>(b, I, scr, smem, pre); break;
    }
    batch_phase();
    // the universal threshold from the level-1 details: one call site for
    // every kind of item, so the selection is inlined once (a called
    // function would keep its saved registers in scratch memory)
    const int64_t moff = dwt ? 0 : n;
    const int64_t mstride = 1;
    const int64_t mcount = dwt ? (I.J == 1 ? n - 1 : n / 2) : n;
    const double t = block_mad<T, NT>(scr, moff, mstride, mcount, smem, pre) * I.factor;
    if (threadIdx.x == 0 && b.thresh_out) b.thresh_out[I.sig] = t;
    switch (kind_filter(I.kind)) {
      case kHAAR: batch_inv<T, kHAAR, NT>(b, I, scr, smem, t); break;
      case kD4: batch_inv<T, kD4, NT>(b, I, scr, smem, t); break;
      case kC6: batch_inv<T, kC6, NT>(b, I, scr, smem, t); break;
      default: batch_inv<T, kLA8, NT>(b, I, scr, smem, t); break;
    }
    batch_phase();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(b.counter + 1, 1) == int(gridDim.x) - 1) {
      atomicExch(b.counter, 0);
      atomicExch(b.counter + 1, 0);
    }
  }
}

// largest MODWT depth of one batch item
constexpr int kBatchMaxModwtJ = 8;

// the first-digit histogram at kBatchPreOff stays clear of what an FP64
// 512-thread item puts in LDS before its selection: the deepest forward's
// rings (LA8, J = 8, chunks of 1024), the DWT's LDS levels and the
// selection's workspace
static_assert((kBatchMaxModwtJ * 1024 + fwd_win_sum(8, kBatchMaxModwtJ)) * 8 <= int(kBatchPreOff),
              "forward rings overlap the first-digit histogram");
static_assert(2 * batch_lds_elems<double>() * 8 <= int(kBatchPreOff),
              "DWT LDS levels overlap the first-digit histogram");
static_assert(block_sel_lds<512>() <= kBatchPreOff,
              "selection workspace overlaps the first-digit histogram");

struct BatchLaunch {
  int NT;
  unsigned grid;
  size_t lds;
};
// launch of the batch kernel of element type T, and its resident
// workgroups per CU for NT threads and lds bytes of dynamic LDS; defined in
// batch_<type>.hip
template <typename T>
int batch_launch(const BatchLaunch& bl, const BatchArgs& a, hipStream_t st);
template <typename T>
int batch_occupancy(int NT, size_t lds, int* blocks);

}  // namespace wh
