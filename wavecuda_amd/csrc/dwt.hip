// dwt.hip -- level-fused periodic DWT, forward and inverse (gfx950).
//
// Math and layout follow the reference's CPU convolution transforms
// (src/haar.cpp:32-93, src/daub4.cpp:5-88, src/c6.cpp:15-111,
// src/la8.cpp:10-115): level j (s = 2^{j-1}) maps the coefficients
// a[m] = y[m s] of the previous level to
//   s_j[m] = sum_k C_k a[(2m+k) mod N],  d_j[m] = sum_k g_k a[(2m+k) mod N]
// stored in place at y[2m s] / y[(2m+1) s] (convolution alignment, not the
// lifting alignment of the reference's GPU code), and the inverse is the
// transpose: a[2m] = sum_q s[m-q] C_{2q} + d[m-q] g_{2q}, a[2m+1] likewise.
//
// Kernel design.  The transform runs as a few passes over compact
// sequences.  A tile pass fuses g <= 5 levels: each workgroup loads a tile
// of TILE coefficients plus the (2^g - 1)(L - 2) halo its cascade needs
// (forward; the inverse needs at most L - 2 per level below the tile),
// computes all g levels in LDS, writes the details to their interleaved
// positions of y (fully coalesced for the first pass) and hands the coarse
// s on as a compact array, so the next pass reads contiguous memory.  Once
// the compact sequence fits one workgroup's LDS a single "whole" kernel
// finishes every remaining level.  Thresholding is fused into the inverse
// as the details are read.
#include <algorithm>
#include <mutex>
#include <vector>

#include "dwt_dev.h"
#include "wh_common.h"
#include "wh_internal.h"

namespace wh {

namespace {

constexpr int kTile = 2048;   // coefficients per tile (level j0 units)
// Passes over at most 2^19 values (the later passes of a long transform,
// or all of a short one) use half-size tiles: twice the workgroups for a
// grid that would leave CUs idle, each with half the dependent work
constexpr int kTileSmall = 1024;
constexpr int64_t kBigTileN = int64_t(1) << 20;
constexpr int kTileTiny = 512;  // passes over at most 2^15 values
constexpr int64_t kSmallTileN = int64_t(1) << 16;
constexpr int kNT = 256;      // threads per tile workgroup
constexpr int kNTW = 1024;    // threads of the whole-sequence workgroup
constexpr int kGmax = 5;      // levels per tile pass
constexpr int kWholeWork = 2048;  // longest sequence a "whole" kernel finishes

__host__ __device__ constexpr int halo_fwd(int levels, int L) { return ((1 << levels) - 1) * (L - 2); }

// Inverse halos: level l's arrays start e_l positions (of level l) below
// the tile, e_l = h_l rounded up to even with h_l = e_{l-1}/2 + L/2 - 1 (the
// inverse of level l+1 reads s_l and d_l that far back); a pair's two
// outputs then never straddle an array start.  e_l <= L - 2.
__host__ __device__ constexpr int inv_evn(int l, int L) {
  int ee = 0;
  for (int i = 1; i <= l; ++i) ee = ((ee / 2 + L / 2 - 1) + 1) & ~1;
  return ee;
}
// The lowest position of the level-j0 sequence an inverse tile pass reads
// below its tile: d_l from index B_l - e_l, at compact position
// base - e_l 2^l + 2^{l-1}
__host__ __device__ constexpr int inv_halo(int L) {
  int h = 0;
  for (int l = 1; l <= kGmax; ++l) {
    const int v = (inv_evn(l, L) << l) - (1 << (l - 1));
    h = v > h ? v : h;
  }
  return h;
}

struct Thr {
  int on, hardness, minlevel, maxlevel;
  double value;
  const double* levels;
};

template <typename T>
__device__ __forceinline__ T thr_level(const Thr& t, int lev0) {
  if (!t.on || lev0 < t.minlevel || lev0 > t.maxlevel) return T(-1);
  return t.levels ? T(t.levels[lev0]) : T(t.value);
}

struct TileArgs {
  const void* src;   // compact input (fwd: s_{j0}; inv: s_{j0+g} at src[i << src_shift])
  void* y;           // interleaved coefficient array
  void* dst;         // fwd: compact s_{j0+g}; inv: compact s_{j0}
  int64_t n;         // full length
  int64_t N;         // compact length at level j0
  int j0, g;
  int final_pass;    // fwd: s_{j0+g} = s_J goes to y
  int write_coarse;  // fwd: write y[q] with ctz >= g (first pass only)
  int src_shift;
  Thr thr;
};

// ------------------------------------------------------------ forward tile
template <typename T, int FILT, int TILE>
__global__ __launch_bounds__(kNT) void dwt_fwd_tile(TileArgs a) {
  constexpr int L = Filter<FILT>::L;
  constexpr int HMAX = halo_fwd(kGmax, L);
  constexpr int SZA = TILE + HMAX;
  // outputs per thread at level 1 (the widest level); each level's s is
  // computed into registers and written back over its input after a
  // barrier, so one buffer serves every level.  The level-1 details stay in
  // the registers of the thread that writes them out (thread t computes
  // d_1[t + k kNT] and writes y positions 2 (t + k kNT) and the next), so
  // LDS holds only the deeper levels' details: 26 KB in FP64, six
  // workgroups per CU.
  constexpr int kPer = (TILE / 2 + halo_fwd(kGmax - 1, L) + kNT - 1) / kNT;
  constexpr int kPer1 = TILE / 2 / kNT;  // level-1 details a thread keeps
  static_assert(kPer1 * kNT == TILE / 2, "whole level-1 pairs per thread");
  using V = typename Vec2<T>::type;
  typedef T E2 __attribute__((ext_vector_type(2)));
  __shared__ __align__(16) T bufA[SZA + 2];
  __shared__ __align__(16) T dbuf[TILE / 2];  // level l >= 2 at TILE/2 - (TILE >> (l-1))

  const int tid = threadIdx.x;
  const int g = a.g;
  const int64_t N = a.N;
  const T* __restrict__ src = static_cast<const T*>(a.src);
  T* __restrict__ y = static_cast<T*>(a.y);

  const int64_t base = int64_t(blockIdx.x) * TILE;
  const int K0 = TILE + halo_fwd(g, L);
  // every load of the tile is issued before the first one is used (a loop
  // that stores each load to LDS before issuing the next waits one memory
  // latency per iteration); only the last tile's halo wraps around
  constexpr int kLd = (SZA / 2 + kNT - 1) / kNT;
  auto stage = [&](auto addr) {
    E2 ld[kLd];  // (an array of HIP's double2 / float2 structs would live in scratch)
#pragma unroll
    for (int k = 0; k < kLd; ++k) {
      const int q = 2 * (tid + k * kNT);
      // unconditional (a clamped duplicate past the end): a guarded load
      // would be merged with its guarded store and wait there
      ld[k] = *reinterpret_cast<const E2*>(addr(q < K0 ? q : K0 - 2));
    }
#pragma unroll
    for (int k = 0; k < kLd; ++k) {
      const int q = 2 * (tid + k * kNT);
      if (q < K0) *reinterpret_cast<E2*>(bufA + q) = ld[k];
    }
  };
  if (base + K0 <= N) {
    const T* sp = src + base;
    stage([&](int q) { return sp + q; });
  } else {
    stage([&](int q) { return src + ((base + q) & (N - 1)); });
  }
  __syncthreads();

  T* cur = bufA;
  T d1[kPer1];
  for (int l = 1; l <= g; ++l) {
    const int own = TILE >> l;
    const int Kl = own + halo_fwd(g - l, L);
    const int doff = TILE / 2 - (TILE >> (l - 1));  // levels >= 2
    T sreg[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * kNT;
      if (i < Kl) {
        // taps 2i .. 2i+L-1 as L/2 aligned 16-byte pairs (no bank conflicts)
        T v[L];
#pragma unroll
        for (int kk = 0; kk < L / 2; ++kk) {
          const V p = *reinterpret_cast<const V*>(cur + 2 * i + 2 * kk);
          v[2 * kk] = p.x;
          v[2 * kk + 1] = p.y;
        }
        T sv, d;
        fwd_pair<T, FILT>(v, sv, d);
        sreg[k] = sv;
        if (l == 1) {
          if (k < kPer1) d1[k < kPer1 ? k : 0] = d;  // i < own exactly when k < kPer1
        } else if (i < own) {
          dbuf[doff + i] = d;
        }
      }
    }
    __syncthreads();  // every tap of this level read before s_l overwrites it
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * kNT;
      if (i < Kl) cur[i] = sreg[k];
    }
    __syncthreads();
  }
  // cur holds s_{j0+g}; own part is cur[0 .. TILE >> g).  Thread t writes
  // the pairs of positions q = 2 (t + k kNT), q + 1: q holds a coarser
  // coefficient (a deeper detail, or s when its ctz is at least g), q + 1
  // the level-1 detail the thread kept.
  const int j0 = a.j0;
  const bool coarse = a.final_pass || a.write_coarse;
#pragma unroll
  for (int k = 0; k < kPer1; ++k) {
    const int qq = tid + k * kNT;
    const int q = 2 * qq;
    const int z = (q == 0) ? 31 : __builtin_ctz(unsigned(q));
    const T v0 = (z >= g) ? cur[q >> g] : dbuf[(TILE / 2 - (TILE >> z)) + (q >> (z + 1))];
    if (j0 == 0 && a.write_coarse) {
      // first pass: every position of the tile, as 16-byte pairs
      *reinterpret_cast<V*>(y + base + q) = V{v0, d1[k]};
    } else {
      if (z < g || coarse) y[(base + q) << j0] = v0;
      y[(base + q + 1) << j0] = d1[k];
    }
  }
  if (!a.final_pass) {
    T* __restrict__ dst = static_cast<T*>(a.dst);
    const int own = TILE >> g;
    for (int i = tid; i < own; i += kNT) dst[(base >> g) + i] = cur[i];
  }
}

// ------------------------------------------------------------ inverse tile
// A workgroup's work is a short chain of dependent phases, so its length
// and the number of workgroups a CU holds, not bandwidth, set the pass's
// time.  The kernel keeps the chain short and its LDS small:
//   A  stage the y span (the level-j0 sequence around the tile) and the top
//      s_{j0+g}, every load issued first;
//   B  copy the details of levels 3..g out of the span into compact arrays
//      (all LDS reads issued first, thresholded on the way);
//   C  levels g..4 write s_{l-1} into two small buffers in turn, level 3
//      writes s_2 into the span itself, at the positions of the coarser
//      coefficients it no longer needs (one barrier a level);
//   D  level 2 runs in place in the span (s_2 and d_2 at positions 4i and
//      4i + 2 of the level-j0 sequence, s_1 written back over them after a
//      barrier), level 1 too: s_1 and d_1 sit side by side there, one
//      16-byte read a tap, and its outputs go straight to memory.
// 25 KB of LDS in FP64 with LA8: six workgroups per CU.
template <typename T, int FILT, int TILE>
__global__ __launch_bounds__(kNT) void dwt_inv_tile(TileArgs a) {
  constexpr int L = Filter<FILT>::L;
  constexpr int PADS = 8;
  // the span every pass stages: the tile and the HY positions below it
  // down to the lowest detail any of its levels reads (an even count)
  constexpr int HY = inv_halo(L);
  static_assert(HY % 2 == 0, "the span starts on a pair");
  constexpr int kSpan = TILE + HY;
  typedef T E2 __attribute__((ext_vector_type(2)));
  // rest: d_3..d_kGmax (level l at doff(l), own_l + e_l values and PADS
  // spare), then P0 (s_4) and P1 (s_3, s_5); the top s_g of a pass with
  // g >= 3 goes where level g reads it (P0 or P1)
  constexpr int kD = (TILE / 4 - (TILE >> kGmax)) + PADS * (kGmax - 2);
  constexpr int kP0 = TILE / 16 + L - 2;  // s_4: own_4 + e_4
  constexpr int kP1 = TILE / 8 + L - 2;   // s_3: own_3 + e_3
  static_assert(kP0 % 2 == 0 && kP1 % 2 == 0 && kD % 2 == 0, "pair-aligned buffers");
  __shared__ __align__(16) T pool[kSpan];
  __shared__ __align__(16) T rest[kD + kP0 + kP1];
  T* P0 = rest + kD;
  T* P1 = P0 + kP0;

  const int tid = threadIdx.x;
  const int g = a.g, j0 = a.j0;
  const int64_t N = a.N;
  const int64_t base = int64_t(blockIdx.x) * TILE;
  const T* __restrict__ src = static_cast<const T*>(a.src);
  const T* __restrict__ y = static_cast<const T*>(a.y);

  auto evn = [](int l) { return inv_evn(l, L); };
  auto doff = [](int l) { return (TILE / 4 - (TILE >> (l - 1))) + PADS * (l - 3); };
  auto sbuf = [&](int j) -> T* { return (j & 1) ? P1 : P0; };  // s_j, j >= 3

  // Local index i of level l's arrays (s_l, d_l) is level-l index
  // m = B_l - e_l + i; s_l[m] sits at compact position m 2^l, d_l[m] at
  // (2m + 1) 2^{l-1}, and the span slot of compact position c is
  // (c - sbase) mod N, so d_l's slot is ((2i + 1) 2^{l-1} + HY - e_l 2^l)
  // mod N (B_l 2^l = base), s_1's is (2i + HY - 2 e_1) mod N and s_2's
  // (4i + HY - 4 e_2) mod N.
  const int64_t sbase = base - HY;  // compact position of pool[0]
  const uint32_t pmask = N > int64_t(0xffffffff) ? 0xffffffffu : uint32_t(N - 1);
  auto dslot = [&](int l, int i) -> uint32_t {
    return ((uint32_t(2 * i + 1) << (l - 1)) + uint32_t(HY - (evn(l) << l))) & pmask;
  };
  const uint32_t c1 = uint32_t(HY - 2 * evn(1));  // s_1[i] at slot 2i + c1, d_1[i] next to it
  const uint32_t c2 = uint32_t(HY - 4 * evn(2));  // s_2[i] at slot 4i + c2, d_2[i] 2 further
  static_assert(HY >= 4 * inv_evn(2, L) && HY >= 2 * inv_evn(1, L), "s_1, s_2 inside the span");

  // ---- A: the span [base - HY, base + TILE).  The first pass (j0 = 0)
  // reads it with contiguous 16-byte loads; deeper passes with one load per
  // value.  Loads past the end are clamped duplicates: a guarded load would
  // be merged with its guarded store and wait there.
  constexpr int kLdP = (kSpan / 2 + kNT - 1) / kNT;  // pair loads a thread (j0 = 0)
  constexpr int kLdS = (kSpan + kNT - 1) / kNT;      // single loads a thread (j0 > 0)
  constexpr int kTopMax = (TILE / 2 + L + kNT - 1) / kNT;
  const int span = int(std::min<int64_t>(kSpan, N));  // pool positions used
  const int own_g = TILE >> g;
  const int eg = evn(g);
  const int cnt_g = own_g + eg;
  const int64_t Bg = base >> g;
  const int64_t Ng = N >> g;
  // top s_{j0+g}: one load a thread unless the pass has fewer than kGmax
  // levels (a uniform branch)
  T topv[kTopMax];
  auto load_top = [&] {
    topv[0] = src[((Bg - eg + (tid < cnt_g ? tid : cnt_g - 1)) & (Ng - 1)) << a.src_shift];
    if (cnt_g > kNT) {
#pragma unroll
      for (int k = 1; k < kTopMax; ++k) {
        const int i = tid + k * kNT;
        const int ic = i < cnt_g ? i : cnt_g - 1;
        topv[k] = src[((Bg - eg + ic) & (Ng - 1)) << a.src_shift];
      }
    }
  };
  // only the first tile's halo (or a sequence shorter than the span) wraps
  // around; every other tile addresses its span from one base pointer
  const bool wrap = (base < HY) || (N < kSpan);
  if (j0 == 0) {
    E2 ld[kLdP];
    if (!wrap) {
      const T* ys = y + sbase;
#pragma unroll
      for (int k = 0; k < kLdP; ++k) {
        const int i = 2 * (tid + k * kNT);
        ld[k] = *reinterpret_cast<const E2*>(ys + (i < span ? i : span - 2));
      }
    } else {
#pragma unroll
      for (int k = 0; k < kLdP; ++k) {
        const int i = 2 * (tid + k * kNT);
        const int ic = i < span ? i : span - 2;
        ld[k] = *reinterpret_cast<const E2*>(y + ((sbase + ic) & (N - 1)));
      }
    }
    load_top();
#pragma unroll
    for (int k = 0; k < kLdP; ++k) {
      const int i = 2 * (tid + k * kNT);
      if (i < span) *reinterpret_cast<E2*>(pool + i) = ld[k];
    }
  } else {
    T ld[kLdS];
    if (!wrap) {
      const T* ys = y + (sbase << j0);
#pragma unroll
      for (int k = 0; k < kLdS; ++k) {
        const int i = tid + k * kNT;
        ld[k] = ys[uint64_t(i < span ? i : span - 1) << j0];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kLdS; ++k) {
        const int i = tid + k * kNT;
        const int ic = i < span ? i : span - 1;
        ld[k] = y[((sbase + ic) & (N - 1)) << j0];
      }
    }
    load_top();
#pragma unroll
    for (int k = 0; k < kLdS; ++k) {
      const int i = tid + k * kNT;
      if (i < span) pool[i] = ld[k];
    }
  }
  if (g >= 3) {
    T* top = sbuf(g);
#pragma unroll
    for (int k = 0; k < kTopMax; ++k) {
      const int i = tid + k * kNT;
      if (i < cnt_g) top[i] = topv[k];
    }
  }
  __syncthreads();
  if (g <= 2) {
    // a one- or two-level pass: its top goes where level g reads it, over
    // the span's coarser positions (staged above, hence the extra barrier)
    const uint32_t cg = g == 1 ? c1 : c2;
#pragma unroll
    for (int k = 0; k < kTopMax; ++k) {
      const int i = tid + k * kNT;
      if (i < cnt_g) pool[((uint32_t(i) << g) + cg) & pmask] = topv[k];
    }
    __syncthreads();
  }

  // ---- B: d_3..d_g out of the span, every read issued first (for levels
  // past g too: their clamped slots stay inside the span)
  if (g >= 3) {
    constexpr int kIt3 = (TILE / 8 + L + kNT - 1) / kNT;
    T ev[kGmax][kIt3];
#pragma unroll
    for (int l = 3; l <= kGmax; ++l) {
      const int cnt = (TILE >> l) + evn(l);
#pragma unroll
      for (int k = 0; k < ((TILE >> l) + L + kNT - 1) / kNT; ++k) {
        const int i = tid + k * kNT;
        ev[l - 1][k] = pool[dslot(l, i < cnt ? i : cnt - 1)];
      }
    }
#pragma unroll
    for (int l = 3; l <= kGmax; ++l) {
      if (l <= g) {
        const int cnt = (TILE >> l) + evn(l);
        const T t = thr_level<T>(a.thr, j0 + l - 1);
        T* dl = rest + doff(l);
#pragma unroll
        for (int k = 0; k < ((TILE >> l) + L + kNT - 1) / kNT; ++k) {
          const int i = tid + k * kNT;
          if (i < cnt) {
            T d = ev[l - 1][k];
            if (t >= T(0)) d = thresh_coef(d, t, a.thr.hardness);
            dl[i] = d;
          }
        }
      }
    }
    __syncthreads();
  }

  // ---- C: levels g..3.  Pair p of level l has local index
  // e_l - e_{l-1}/2 + p in s_l and d_l, and outputs local 2p, 2p + 1 of
  // s_{l-1}.  Every pair is computed before any is written: the compiler
  // cannot tell the buffers apart, so a write would hold back later reads.
  constexpr int kPairs = (TILE / 2 + kNT - 1) / kNT;
  for (int l = g; l >= 3; --l) {
    const T* cur = sbuf(l);
    const T* dl = rest + doff(l);
    const int elm = evn(l - 1);
    const int npairs = ((TILE >> (l - 1)) + elm) / 2;
    const int off = evn(l) - elm / 2;
    E2 res[kPairs];
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int p = tid + k * kNT;
      if (p < npairs) {
        T sv[L / 2], dv[L / 2];
#pragma unroll
        for (int q = 0; q < L / 2; ++q) {
          sv[q] = cur[off + p - q];
          dv[q] = dl[off + p - q];
        }
        T e, o;
        inv_pair<T, FILT>(sv, dv, e, o);
        res[k] = E2{e, o};
      }
    }
    if (l > 3) {
      T* nxt = sbuf(l - 1);
#pragma unroll
      for (int k = 0; k < kPairs; ++k) {
        const int p = tid + k * kNT;
        if (p < npairs) *reinterpret_cast<E2*>(nxt + 2 * p) = res[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kPairs; ++k) {
        const int p = tid + k * kNT;
        if (p < npairs) {
          pool[(uint32_t(8 * p) + c2) & pmask] = res[k].x;
          pool[(uint32_t(8 * p + 4) + c2) & pmask] = res[k].y;
        }
      }
    }
    __syncthreads();
  }

  // ---- level 2 in place: read every tap, barrier, write s_1 over them
  if (g >= 2) {
    const int elm = evn(1);
    const int npairs = ((TILE >> 1) + elm) / 2;
    const int off = evn(2) - elm / 2;
    const T t = thr_level<T>(a.thr, j0 + 1);
    E2 res[kPairs];
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int p = tid + k * kNT;
      if (p < npairs) {
        T sv[L / 2], dv[L / 2];
#pragma unroll
        for (int q = 0; q < L / 2; ++q) {
          const uint32_t sl = uint32_t(4 * (off + p - q)) + c2;
          sv[q] = pool[sl & pmask];
          const T d = pool[(sl + 2) & pmask];
          dv[q] = (t >= T(0)) ? thresh_coef(d, t, a.thr.hardness) : d;
        }
        T e, o;
        inv_pair<T, FILT>(sv, dv, e, o);
        res[k] = E2{e, o};
      }
    }
    __syncthreads();  // every tap read before s_1 overwrites them
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int p = tid + k * kNT;
      if (p < npairs) {
        pool[(uint32_t(4 * p) + c1) & pmask] = res[k].x;
        pool[(uint32_t(4 * p + 2) + c1) & pmask] = res[k].y;
      }
    }
    __syncthreads();
  }

  // ---- D: level 1 in place in the span; outputs x[base + 2p], x[base + 2p + 1]
  {
    const int npairs = TILE / 2;  // e_0 = 0
    const int off = evn(1);
    const T t = thr_level<T>(a.thr, j0);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    E2 res[kPairs];
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int p = tid + k * kNT;
      if (p < npairs) {
        T sv[L / 2], dv[L / 2];
#pragma unroll
        for (int q = 0; q < L / 2; ++q) {
          const E2 v = *reinterpret_cast<const E2*>(pool + ((uint32_t(2 * (off + p - q)) + c1) & pmask));
          sv[q] = v.x;
          dv[q] = (t >= T(0)) ? thresh_coef(v.y, t, a.thr.hardness) : v.y;
        }
        T e, o;
        inv_pair<T, FILT>(sv, dv, e, o);
        res[k] = E2{e, o};
      }
    }
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int p = tid + k * kNT;
      if (p < npairs) *reinterpret_cast<E2*>(dst + base + 2 * p) = res[k];
    }
  }
}

// ----------------------------------------------------------- whole kernels
struct WholeArgs {
  const void* src;  // fwd: compact s_{j0} (N values); inv: y
  void* y;
  void* dst;        // inv: compact s_{j0}
  int64_t N;        // compact length at level j0
  int j0, J;
  Thr thr;
};

template <typename T, int FILT>
__global__ __launch_bounds__(kNTW) void dwt_fwd_whole(WholeArgs a) {
  constexpr int L = Filter<FILT>::L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + a.N;
  const int tid = threadIdx.x;
  const T* __restrict__ src = static_cast<const T*>(a.src);
  T* __restrict__ y = static_cast<T*>(a.y);
  const int N = int(a.N);
  for (int i = tid; i < N; i += kNTW) A[i] = src[i];
  __syncthreads();
  T* cur = A;
  T* nxt = B;
  int Nl = N;
  for (int j = a.j0 + 1; j <= a.J; ++j) {
    const int P = Nl / 2;
    for (int m = tid; m < P; m += kNTW) {
      T v[L];
#pragma unroll
      for (int k = 0; k < L; ++k) v[k] = cur[(2 * m + k) & (Nl - 1)];
      T s, d;
      fwd_pair<T, FILT>(v, s, d);
      nxt[m] = s;
      y[int64_t(2 * m + 1) << (j - 1)] = d;
    }
    __syncthreads();
    T* t = cur;
    cur = nxt;
    nxt = t;
    Nl = P;
  }
  for (int m = tid; m < Nl; m += kNTW) y[int64_t(m) << a.J] = cur[m];
}

template <typename T, int FILT>
__global__ __launch_bounds__(kNTW) void dwt_inv_whole(WholeArgs a) {
  constexpr int L = Filter<FILT>::L;
  constexpr int kLd = (kWholeWork + kNTW - 1) / kNTW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* B = A + a.N;
  T* Cf = B + a.N / 2;  // the N coefficients of levels j0+1..J, thresholded
  const int tid = threadIdx.x;
  const T* __restrict__ y = static_cast<const T*>(a.src);
  T* __restrict__ dst = static_cast<T*>(a.dst);
  const int N = int(a.N);
  const int j0 = a.j0, J = a.J;
  // The level-j0 sequence y[q 2^j0], q < N, holds every coefficient the
  // remaining levels read: s_J where q is a multiple of 2^{J-j0}, else the
  // detail of level j0 + ctz(q) + 1.  It is loaded once, every load issued
  // before the first is used, and thresholded as it goes to LDS.
  T ld[kLd];
#pragma unroll
  for (int k = 0; k < kLd; ++k) {
    const int q = tid + k * kNTW;
    const int qc = q < N ? q : N - 1;  // unconditional load, clamped
    ld[k] = y[int64_t(qc) << j0];
  }
#pragma unroll
  for (int k = 0; k < kLd; ++k) {
    const int q = tid + k * kNTW;
    if (q < N) {
      const int z = (q == 0) ? 31 : __builtin_ctz(unsigned(q));
      T d = ld[k];
      if (z < J - j0) {
        const T t = thr_level<T>(a.thr, j0 + z);
        if (t >= T(0)) d = thresh_coef(d, t, a.thr.hardness);
      }
      Cf[q] = d;
    }
  }
  __syncthreads();
  // s_J into the buffer that level J reads: output of level j goes to A
  // when (j - j0) is odd, B otherwise; level J reads the output slot of J+1
  int NJ = N >> (J - j0);
  T* top = ((J + 1 - j0) & 1) ? A : B;
  for (int m = tid; m < NJ; m += kNTW) top[m] = Cf[m << (J - j0)];
  __syncthreads();
  int Nl = NJ;  // length of s_j
  for (int j = J; j > j0; --j) {
    const T* cur = ((j + 1 - j0) & 1) ? A : B;
    T* out = ((j - j0) & 1) ? A : B;
    const int sh = j - 1 - j0;
    for (int m = tid; m < Nl; m += kNTW) {
      T sv[L / 2], dv[L / 2];
#pragma unroll
      for (int q = 0; q < L / 2; ++q) {
        int mm = (m - q) & (Nl - 1);
        sv[q] = cur[mm];
        dv[q] = Cf[(2 * mm + 1) << sh];
      }
      T ev, ov;
      inv_pair<T, FILT>(sv, dv, ev, ov);
      out[2 * m] = ev;
      out[2 * m + 1] = ov;
    }
    __syncthreads();
    Nl *= 2;
  }
  // the output of level j0+1 is in A
  for (int i = tid; i < N; i += kNTW) dst[i] = A[i];
}

// ------------------------------------------------------------------- host
struct Pass {
  bool whole;
  int j0, g;
  int64_t N;
};

// Largest compact length finished by one workgroup: its LDS (2.5 N values
// for the inverse) must fit the budget, and the one workgroup's serial walk
// must stay short (kWholeWork, with kTile: a tile pass of one or two
// workgroups is cheaper than one workgroup over more values).

template <typename T>
int64_t whole_max() {
  const size_t budget = std::min<size_t>(device_info().max_lds_per_block, 160 * 1024);
  int64_t m = 1;
  while (m * 2 <= kWholeWork && size_t(5 * m) * sizeof(T) <= budget) m *= 2;
  return m;
}

template <typename KERN>
void allow_big_lds(KERN kern) {
  allow_dynamic_lds(reinterpret_cast<const void*>(kern));
}

template <typename T>
std::vector<Pass> plan_passes(int64_t n, int J) {
  std::vector<Pass> ps;
  int j0 = 0;
  int64_t N = n;
  while (j0 < J) {
    if (N <= whole_max<T>() || N < int64_t(kTile)) {
      ps.push_back({true, j0, J - j0, N});
      break;
    }
    int g = std::min(kGmax, J - j0);
    ps.push_back({false, j0, g, N});
    j0 += g;
    N >>= g;
  }
  return ps;
}

Thr thr_of(const wh_thresh* t) {
  Thr r{};
  if (t && t->enabled) {
    r.on = 1;
    r.hardness = t->hardness;
    r.minlevel = t->minlevel;
    r.maxlevel = t->maxlevel;
    r.value = t->value;
    r.levels = t->levels;
  }
  return r;
}

template <typename T, int FILT>
int dwt_fwd_f(const void* x, void* y, int64_t n, int J, hipStream_t st) {
  auto ps = plan_passes<T>(n, J);
  const T* src = static_cast<const T*>(x);
  // in place: the first pass reads halos of neighbouring tiles, so stage x
  if (x == y && !ps.front().whole) {
    void* p = nullptr;
    int rc = scratch2_get(size_t(n) * sizeof(T), &p);
    if (rc) return rc;
    WH_HIP_CHECK(hipMemcpyAsync(p, x, size_t(n) * sizeof(T), hipMemcpyDeviceToDevice, st));
    src = static_cast<const T*>(p);
  }
  T* comp = nullptr;
  if (ps.size() > 1) {
    void* p = nullptr;
    int rc = scratch_get(size_t(n >> kGmax) * 2 * sizeof(T) + 64, &p);
    if (rc) return rc;
    comp = static_cast<T*>(p);
  }
  int slot = 0;
  for (size_t pi = 0; pi < ps.size(); ++pi) {
    const Pass& P = ps[pi];
    if (P.whole) {
      WholeArgs a{};
      a.src = src;
      a.y = y;
      a.N = P.N;
      a.j0 = P.j0;
      a.J = J;
      size_t lds = size_t(P.N + P.N / 2) * sizeof(T);
      allow_big_lds(dwt_fwd_whole<T, FILT>);
      hipLaunchKernelGGL((dwt_fwd_whole<T, FILT>), dim3(1), dim3(kNTW), lds, st, a);
      WH_HIP_CHECK(hipGetLastError());
    } else {
      TileArgs a{};
      a.src = src;
      a.y = y;
      T* out = comp + slot * (n >> kGmax);
      a.dst = out;
      a.n = n;
      a.N = P.N;
      a.j0 = P.j0;
      a.g = P.g;
      a.final_pass = (P.j0 + P.g == J);
      a.write_coarse = (P.j0 == 0);
      if (P.N >= kBigTileN)
        hipLaunchKernelGGL((dwt_fwd_tile<T, FILT, kTile>), dim3(unsigned(P.N / kTile)), dim3(kNT),
                           0, st, a);
      else if (P.N >= kSmallTileN)
        hipLaunchKernelGGL((dwt_fwd_tile<T, FILT, kTileSmall>), dim3(unsigned(P.N / kTileSmall)),
                           dim3(kNT), 0, st, a);
      else
        hipLaunchKernelGGL((dwt_fwd_tile<T, FILT, kTileTiny>), dim3(unsigned(P.N / kTileTiny)),
                           dim3(kNT), 0, st, a);
      WH_HIP_CHECK(hipGetLastError());
      src = out;
      slot ^= 1;
    }
  }
  return WH_OK;
}

template <typename T, int FILT>
int dwt_inv_f(const void* yin, void* x, int64_t n, int J, const wh_thresh* thr,
              hipStream_t st) {
  auto ps = plan_passes<T>(n, J);
  const T* y = static_cast<const T*>(yin);
  if (yin == x && !ps.front().whole) {
    void* p = nullptr;
    int rc = scratch2_get(size_t(n) * sizeof(T), &p);
    if (rc) return rc;
    WH_HIP_CHECK(hipMemcpyAsync(p, yin, size_t(n) * sizeof(T), hipMemcpyDeviceToDevice, st));
    y = static_cast<const T*>(p);
  }
  T* comp = nullptr;
  if (ps.size() > 1) {
    void* p = nullptr;
    int rc = scratch_get(size_t(n >> kGmax) * 2 * sizeof(T) + 64, &p);
    if (rc) return rc;
    comp = static_cast<T*>(p);
  }
  const Thr th = thr_of(thr);
  const T* src = nullptr;  // compact s handed down between passes
  int slot = 0;
  for (int pi = int(ps.size()) - 1; pi >= 0; --pi) {
    const Pass& P = ps[pi];
    T* out = (P.j0 == 0) ? static_cast<T*>(x) : comp + slot * (n >> kGmax);
    if (P.whole) {
      WholeArgs a{};
      a.src = y;
      a.dst = out;
      a.N = P.N;
      a.j0 = P.j0;
      a.J = J;
      a.thr = th;
      size_t lds = size_t(2 * P.N + P.N / 2) * sizeof(T);
      allow_big_lds(dwt_inv_whole<T, FILT>);
      hipLaunchKernelGGL((dwt_inv_whole<T, FILT>), dim3(1), dim3(kNTW), lds, st, a);
      WH_HIP_CHECK(hipGetLastError());
    } else {
      TileArgs a{};
      const bool top = (P.j0 + P.g == J);
      a.src = top ? static_cast<const void*>(y) : static_cast<const void*>(src);
      a.src_shift = top ? J : 0;
      a.y = const_cast<T*>(y);
      a.dst = out;
      a.n = n;
      a.N = P.N;
      a.j0 = P.j0;
      a.g = P.g;
      a.thr = th;
      if (P.N >= kBigTileN)
        hipLaunchKernelGGL((dwt_inv_tile<T, FILT, kTile>), dim3(unsigned(P.N / kTile)), dim3(kNT),
                           0, st, a);
      else if (P.N >= kSmallTileN)
        hipLaunchKernelGGL((dwt_inv_tile<T, FILT, kTileSmall>), dim3(unsigned(P.N / kTileSmall)),
                           dim3(kNT), 0, st, a);
      else
        hipLaunchKernelGGL((dwt_inv_tile<T, FILT, kTileTiny>), dim3(unsigned(P.N / kTileTiny)),
                           dim3(kNT), 0, st, a);
      WH_HIP_CHECK(hipGetLastError());
    }
    src = out;
    slot ^= 1;
  }
  return WH_OK;
}

template <typename T>
int dwt_fwd_t(const void* x, void* y, int64_t n, int J, int filt, hipStream_t st) {
  switch (filt) {
    case kHAAR: return dwt_fwd_f<T, kHAAR>(x, y, n, J, st);
    case kD4: return dwt_fwd_f<T, kD4>(x, y, n, J, st);
    case kC6: return dwt_fwd_f<T, kC6>(x, y, n, J, st);
    default: return dwt_fwd_f<T, kLA8>(x, y, n, J, st);
  }
}

template <typename T>
int dwt_inv_t(const void* y, void* x, int64_t n, int J, int filt, const wh_thresh* thr,
              hipStream_t st) {
  switch (filt) {
    case kHAAR: return dwt_inv_f<T, kHAAR>(y, x, n, J, thr, st);
    case kD4: return dwt_inv_f<T, kD4>(y, x, n, J, thr, st);
    case kC6: return dwt_inv_f<T, kC6>(y, x, n, J, thr, st);
    default: return dwt_inv_f<T, kLA8>(y, x, n, J, thr, st);
  }
}

}  // namespace

int dwt_forward(const void* x, void* y, int64_t n, int J, int filt, int dtype, hipStream_t st) {
  return dtype == WH_F32 ? dwt_fwd_t<float>(x, y, n, J, filt, st)
                         : dwt_fwd_t<double>(x, y, n, J, filt, st);
}

int dwt_inverse(const void* y, void* x, int64_t n, int J, int filt, int dtype,
                const wh_thresh* thr, hipStream_t st) {
  return dtype == WH_F32 ? dwt_inv_t<float>(y, x, n, J, filt, thr, st)
                         : dwt_inv_t<double>(y, x, n, J, filt, thr, st);
}

}  // namespace wh
