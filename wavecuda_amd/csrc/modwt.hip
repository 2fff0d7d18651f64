// modwt.hip -- level-fused time-ordered MODWT, forward and inverse (gfx950).
//
// Math (reference src/daub4.cpp:561-589, src/haar.cpp:590-616; C6/LA8
// use the same pattern): with s = 2^{j-1}
//   S_j[i] = sum_k C_k S_{j-1}[(i + k s) mod n],  D_j[i] = sum_k g_k S_{j-1}[...]
//   S_{j-1}[i] = 1/2 sum_k (C_k S_j[(i - k s) mod n] + g_k D_j[(i - k s) mod n])
// stored as xmod[2n(j-1) + i] = S_j[i], xmod[2n(j-1) + n + i] = D_j[i].
//
// Kernel design.  A "group" fuses g consecutive levels j0+1..j0+g.  Level
// j0+l only couples positions congruent modulo B = 2^{j0}, so position
// i = r + B m is split into a residue r and an index m; inside the group
// level l has stride 2^{l-1} in m.  A workgroup owns R residues (R = 1 for
// the group that starts at level 1, R = 32 otherwise: 256-byte HBM runs)
// and one segment [mlo, mhi) of m, and marches through it in chunks of M
// (descending for the forward, whose taps look ahead; ascending for the
// inverse, whose taps look back).  Every level keeps an LDS ring with the
// chunk plus the (L-1) 2^{l-1} window values left from earlier chunks, so
// all levels of a group read their input from LDS and HBM sees each group
// input once and each output once.  A segment is entered after a warm-up of
// (L-1)(2^g - 1) positions (rounded up to M) whose outputs are computed but
// not stored -- the only redundant work.
//
// Each thread owns two adjacent elements (consecutive m when R = 1,
// consecutive residues otherwise), so LDS reads/writes and HBM loads/stores
// are 16-byte vectors; ring slots are kept even so a pair never straddles
// the ring's wrap.  Levels are unrolled with compile-time strides, the next
// chunk's HBM loads are issued before a chunk's arithmetic and parked in
// LDS after it, and the LDS layout [slot][residue] keeps lanes contiguous.
// The threshold of GPUSmooth is applied to the detail blocks as the
// inverse loads them.
#include <algorithm>
#include <mutex>
#include <vector>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {

struct ModwtArgs {
  const void* in;        // fwd: group input S_{j0} (x when j0 == 0)
  void* xmod;            // fwd: output blocks; inv: detail blocks
  const void* stop;      // inv: S_{j0+g} (xmod slot or scratch)
  void* out;             // inv: S_{j0} destination (x when j0 == 0)
  int64_t n;
  int64_t P;             // segment length in m units
  int j0, g, M;
  int warm;              // warm-up in m units, multiple of M
  int nchunks;           // (P + warm) / M
  // fused threshold (inverse only)
  int th_on, hardness, minlevel, maxlevel;
  double th_value;
  const double* th_levels;
};

constexpr int kGmax = 10;

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

__host__ __device__ constexpr int even_up(int v) { return (v + 1) & ~1; }

// Ring geometry in m units.  Forward ring l (l = 0..g-1) holds the input of
// level l+1, window (L-1) 2^l; inverse ring l (l = 1..g) holds S_l and D_l,
// window (L-1) 2^{l-1}.  Windows are rounded up to even.
__host__ __device__ constexpr int fwd_win(int L, int l) { return even_up((L - 1) << l); }
__host__ __device__ constexpr int inv_win(int L, int l) { return even_up((L - 1) << (l - 1)); }
__host__ __device__ inline int fwd_off(int L, int M, int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += M + fwd_win(L, i);
  return o;
}
__host__ __device__ inline int inv_off(int L, int M, int l) {
  int o = 0;
  for (int i = 1; i < l; ++i) o += M + inv_win(L, i);
  return o;
}

template <typename T>
__device__ __forceinline__ typename Vec2<T>::type ldv(const T* p) {
  return *reinterpret_cast<const typename Vec2<T>::type*>(p);
}
template <typename T>
__device__ __forceinline__ void stv(T* p, typename Vec2<T>::type v) {
  *reinterpret_cast<typename Vec2<T>::type*>(p) = v;
}
template <typename T>
__device__ __forceinline__ void stv_nt(T* p, typename Vec2<T>::type v) {
  __builtin_nontemporal_store(v.x, p);
  __builtin_nontemporal_store(v.y, p + 1);
}

// ---------------------------------------------------------------- forward
template <typename T, int FILT, int R, int NT>
struct FwdCtx {
  using V = typename Vec2<T>::type;
  static constexpr int L = Filter<FILT>::L;
  T* lds;
  T* xmod;
  int M, mm, r, g, j0;
  bool active;
  int64_t n, r0, mlo, mhi, c;
  int head[kGmax];  // ring heads (slot of position c), ring l = 0..g-1

  // level l (1-based within the group): ring l-1 -> S_{j0+l}, D_{j0+l}
  template <int l>
  __device__ __forceinline__ void level() {
    constexpr int s = 1 << (l - 1);
    const int Qin = M + fwd_win(L, l - 1);
    const T* rin = lds + size_t(fwd_off(L, M, l - 1)) * R;
    if (active) {
      T s0, d0, s1, d1;
      if constexpr (R == 1 && s == 1) {
        // taps m..m+L-1 and m+1..m+L: L/2+1 aligned pairs from slot m
        T w[L + 2];
        int sl = head[l - 1] + mm;
        if (sl >= Qin) sl -= Qin;
#pragma unroll
        for (int q = 0; q <= L / 2; ++q) {
          int sq = sl + 2 * q;
          if (sq >= Qin) sq -= Qin;
          V v = ldv(rin + sq);
          w[2 * q] = v.x;
          w[2 * q + 1] = v.y;
        }
        if constexpr (FILT == kHAAR) {
          s0 = (w[0] + w[1]) * T(kInvR2);
          d0 = (w[0] - w[1]) * T(kInvR2);
          s1 = (w[1] + w[2]) * T(kInvR2);
          d1 = (w[1] - w[2]) * T(kInvR2);
        } else {
          s0 = w[0] * T(Filter<FILT>::c(0));
          d0 = w[0] * T(gcoef<FILT>(0));
          s1 = w[1] * T(Filter<FILT>::c(0));
          d1 = w[1] * T(gcoef<FILT>(0));
#pragma unroll
          for (int k = 1; k < L; ++k) {
            s0 = fma(w[k], T(Filter<FILT>::c(k)), s0);
            d0 = fma(w[k], T(gcoef<FILT>(k)), d0);
            s1 = fma(w[k + 1], T(Filter<FILT>::c(k)), s1);
            d1 = fma(w[k + 1], T(gcoef<FILT>(k)), d1);
          }
        }
      } else {
        V v[L];
        int sl = head[l - 1] + mm;
        if (sl >= Qin) sl -= Qin;
#pragma unroll
        for (int k = 0; k < L; ++k) {
          int sk = sl + k * s;
          if (sk >= Qin) sk -= Qin;
          v[k] = ldv(rin + size_t(sk) * R + r);
        }
        if constexpr (FILT == kHAAR) {
          s0 = (v[0].x + v[1].x) * T(kInvR2);
          d0 = (v[0].x - v[1].x) * T(kInvR2);
          s1 = (v[0].y + v[1].y) * T(kInvR2);
          d1 = (v[0].y - v[1].y) * T(kInvR2);
        } else {
          s0 = v[0].x * T(Filter<FILT>::c(0));
          d0 = v[0].x * T(gcoef<FILT>(0));
          s1 = v[0].y * T(Filter<FILT>::c(0));
          d1 = v[0].y * T(gcoef<FILT>(0));
#pragma unroll
          for (int k = 1; k < L; ++k) {
            s0 = fma(v[k].x, T(Filter<FILT>::c(k)), s0);
            d0 = fma(v[k].x, T(gcoef<FILT>(k)), d0);
            s1 = fma(v[k].y, T(Filter<FILT>::c(k)), s1);
            d1 = fma(v[k].y, T(gcoef<FILT>(k)), d1);
          }
        }
      }
      if (l < g) {
        const int Qo = M + fwd_win(L, l);
        int so = head[l] + mm;
        if (so >= Qo) so -= Qo;
        stv(lds + size_t(fwd_off(L, M, l)) * R + size_t(so) * R + r, V{s0, s1});
      }
      const int64_t m = c + mm;
      if (m >= mlo && m < mhi) {
        const int64_t i = (R == 1) ? m : r0 + r + (m << j0);
        T* oS = xmod + 2 * n * int64_t(j0 + l - 1) + i;
        stv_nt(oS, V{s0, s1});
        stv_nt(oS + n, V{d0, d1});
      }
    }
    __syncthreads();
  }

  template <int l>
  __device__ __forceinline__ void levels_from() {
    if constexpr (l <= kGmax) {
      if (l <= g) {
        level<l>();
        levels_from<l + 1>();
      }
    }
  }
};

template <typename T, int FILT, int R, int NT>
__global__ __launch_bounds__(NT) void modwt_fwd_v2(ModwtArgs a) {
  using V = typename Vec2<T>::type;
  constexpr int L = Filter<FILT>::L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  FwdCtx<T, FILT, R, NT> C;
  C.lds = reinterpret_cast<T*>(smem_raw);
  C.xmod = static_cast<T*>(a.xmod);
  C.M = a.M;
  C.g = a.g;
  C.j0 = a.j0;
  C.n = a.n;
  const int e = 2 * int(threadIdx.x);
  C.active = e < a.M * R;
  C.mm = e / R;
  C.r = e % R;
  const int64_t Nm = a.n >> a.j0;
  const int64_t nrb = (int64_t(1) << a.j0) / R;
  const int64_t rb = int64_t(blockIdx.x) % nrb, seg = int64_t(blockIdx.x) / nrb;
  C.r0 = rb * R;
  C.mlo = seg * a.P;
  C.mhi = C.mlo + a.P;
  const T* __restrict__ in = static_cast<const T*>(a.in);
  const int64_t ctop = C.mhi + a.warm - a.M;
#pragma unroll
  for (int l = 0; l < kGmax; ++l) C.head[l] = 0;

  auto in_index = [&](int64_t m) -> int64_t {
    m &= (Nm - 1);
    return (R == 1) ? m : C.r0 + C.r + (m << a.j0);
  };
  // prologue: chunk 0 into ring 0
  if (C.active) stv(C.lds + size_t(C.mm) * R + C.r, ldv(in + in_index(ctop + C.mm)));
  __syncthreads();

  for (int k = 0; k < a.nchunks; ++k) {
    C.c = ctop - int64_t(k) * a.M;
    const bool have_next = (k + 1 < a.nchunks);
    V pre{};
    if (have_next && C.active) pre = ldv(in + in_index(C.c - a.M + C.mm));
    C.template levels_from<1>();
    // advance the rings by one chunk (descending march: head -= M mod Q)
#pragma unroll
    for (int l = 0; l < kGmax; ++l) {
      const int Q = a.M + fwd_win(L, l);
      int h = C.head[l] - a.M;
      while (h < 0) h += Q;
      C.head[l] = h;
    }
    if (have_next) {
      if (C.active) {
        const int Q0 = a.M + fwd_win(L, 0);
        int so = C.head[0] + C.mm;
        if (so >= Q0) so -= Q0;
        stv(C.lds + size_t(so) * R + C.r, pre);
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- inverse
template <typename T, int FILT, int R, int NT>
struct InvCtx {
  using V = typename Vec2<T>::type;
  static constexpr int L = Filter<FILT>::L;
  T* lds;
  int baseD;
  const ModwtArgs* a;
  const T* xmod;
  const T* stop;
  T* out;
  int M, mm, r, g, j0;
  bool active;
  int64_t n, Nm, r0, mlo, mhi, c;
  int head[kGmax + 1];  // head of ring l (1..g): slot of position c - W'_l
  V pre[kGmax + 1];     // next chunk: D_l (l = 1..g) and S_g (index 0)

  __device__ __forceinline__ int64_t gidx(int64_t m) const {
    m &= (Nm - 1);
    return (R == 1) ? m : r0 + r + (m << j0);
  }
  __device__ __forceinline__ T thr(int l) const {
    const int lev = j0 + l - 1;
    if (!a->th_on || lev < a->minlevel || lev > a->maxlevel) return T(-1);
    return a->th_levels ? T(a->th_levels[lev]) : T(a->th_value);
  }
  __device__ __forceinline__ V thr2(V d, T t) const {
    if (t >= T(0)) {
      d.x = thresh_coef(d.x, t, a->hardness);
      d.y = thresh_coef(d.y, t, a->hardness);
    }
    return d;
  }

  // issue the HBM loads of chunk cc for ring l (l = 1..g)
  template <int l>
  __device__ __forceinline__ void load_next(int64_t cc) {
    if constexpr (l <= kGmax) {
      if (l <= g) {
        if (active) {
          const int64_t i = gidx(cc + mm);
          pre[l] = ldv(xmod + 2 * n * int64_t(j0 + l - 1) + n + i);
          if (l == g) pre[0] = ldv(stop + i);
        }
        load_next<l + 1>(cc);
      }
    }
  }
  // park the loaded chunk in the rings (slots head_l + W'_l + mm)
  template <int l>
  __device__ __forceinline__ void park() {
    if constexpr (l <= kGmax) {
      if (l <= g) {
        if (active) {
          const int Q = M + inv_win(L, l);
          int so = head[l] + inv_win(L, l) + mm;
          if (so >= Q) so -= Q;
          const size_t off = size_t(inv_off(L, M, l)) * R + size_t(so) * R + r;
          stv(lds + baseD + off, thr2(pre[l], thr(l)));
          if (l == g) stv(lds + off, pre[0]);
        }
        park<l + 1>();
      }
    }
  }

  // level l: S_l, D_l (ring l) -> S_{l-1} (ring l-1, or HBM when l == 1)
  template <int l>
  __device__ __forceinline__ void level() {
    constexpr int s = 1 << (l - 1);
    const int W = inv_win(L, l);
    const int Q = M + W;
    const T* rs = lds + size_t(inv_off(L, M, l)) * R;
    const T* rd = rs + baseD;
    if (active) {
      T e0, e1;
      // tap k of element m sits at slot head + W + mm - k s
      if constexpr (R == 1 && s == 1) {
        // pairs from slot head + mm (even): w[i] = slot head + mm + i
        T ws[L + 2], wd[L + 2];
        int sl = head[l] + mm;
        if (sl >= Q) sl -= Q;
#pragma unroll
        for (int q = 0; q <= L / 2; ++q) {
          int sq = sl + 2 * q;
          if (sq >= Q) sq -= Q;
          V vs = ldv(rs + sq), vd = ldv(rd + sq);
          ws[2 * q] = vs.x;
          ws[2 * q + 1] = vs.y;
          wd[2 * q] = vd.x;
          wd[2 * q + 1] = vd.y;
        }
        // element m: tap k at w[L - k]; element m+1 at w[L + 1 - k]
        // (the level-1 window is L - 1 rounded up to L)
        if constexpr (FILT == kHAAR) {
          e0 = (ws[L] + wd[L] + ws[L - 1] - wd[L - 1]) * T(kHInvR2);
          e1 = (ws[L + 1] + wd[L + 1] + ws[L] - wd[L]) * T(kHInvR2);
        } else {
          bool first = true;
          e0 = T(0);
          e1 = T(0);
#pragma unroll
          for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int kk = L - 2 + pass; kk >= 0; kk -= 2) {
              const T cc = T(Filter<FILT>::c(kk)), gk = T(gcoef<FILT>(kk));
              if (first) {
                e0 = ws[L - kk] * cc;
                e1 = ws[L + 1 - kk] * cc;
                first = false;
              } else {
                e0 = fma(ws[L - kk], cc, e0);
                e1 = fma(ws[L + 1 - kk], cc, e1);
              }
              e0 = fma(wd[L - kk], gk, e0);
              e1 = fma(wd[L + 1 - kk], gk, e1);
            }
          }
          e0 = T(0.5) * e0;
          e1 = T(0.5) * e1;
        }
      } else {
        int sl = head[l] + W + mm;
        if (sl >= Q) sl -= Q;
        if constexpr (FILT == kHAAR) {
          int s1 = sl - s;
          if (s1 < 0) s1 += Q;
          V S0 = ldv(rs + size_t(sl) * R + r), D0 = ldv(rd + size_t(sl) * R + r);
          V S1 = ldv(rs + size_t(s1) * R + r), D1 = ldv(rd + size_t(s1) * R + r);
          e0 = (S0.x + D0.x + S1.x - D1.x) * T(kHInvR2);
          e1 = (S0.y + D0.y + S1.y - D1.y) * T(kHInvR2);
        } else {
          V vs[L], vd[L];
#pragma unroll
          for (int kk = 0; kk < L; ++kk) {
            int sk = sl - kk * s;
            if (sk < 0) sk += Q;
            vs[kk] = ldv(rs + size_t(sk) * R + r);
            vd[kk] = ldv(rd + size_t(sk) * R + r);
          }
          bool first = true;
          e0 = T(0);
          e1 = T(0);
          // term order k = L-2, L-4, ..., 0, L-1, ..., 1 (src/daub4.cpp:578-583)
#pragma unroll
          for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int kk = L - 2 + pass; kk >= 0; kk -= 2) {
              const T cc = T(Filter<FILT>::c(kk)), gk = T(gcoef<FILT>(kk));
              if (first) {
                e0 = vs[kk].x * cc;
                e1 = vs[kk].y * cc;
                first = false;
              } else {
                e0 = fma(vs[kk].x, cc, e0);
                e1 = fma(vs[kk].y, cc, e1);
              }
              e0 = fma(vd[kk].x, gk, e0);
              e1 = fma(vd[kk].y, gk, e1);
            }
          }
          e0 = T(0.5) * e0;
          e1 = T(0.5) * e1;
        }
      }
      if (l > 1) {
        const int Wo = inv_win(L, l - 1);
        const int Qo = M + Wo;
        int so = head[l - 1] + Wo + mm;
        if (so >= Qo) so -= Qo;
        stv(lds + size_t(inv_off(L, M, l - 1)) * R + size_t(so) * R + r, V{e0, e1});
      } else {
        const int64_t m = c + mm;
        if (m >= mlo && m < mhi) {
          const int64_t i = (R == 1) ? m : r0 + r + (m << j0);
          stv_nt(out + i, V{e0, e1});
        }
      }
    }
    __syncthreads();
  }

  template <int l>
  __device__ __forceinline__ void levels_down() {
    if constexpr (l >= 1) {
      if (l <= g) level<l>();
      levels_down<l - 1>();
    }
  }
};

template <typename T, int FILT, int R, int NT>
__global__ __launch_bounds__(NT) void modwt_inv_v2(ModwtArgs a) {
  constexpr int L = Filter<FILT>::L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  InvCtx<T, FILT, R, NT> C;
  C.lds = reinterpret_cast<T*>(smem_raw);
  C.a = &a;
  C.xmod = static_cast<const T*>(a.xmod);
  C.stop = static_cast<const T*>(a.stop);
  C.out = static_cast<T*>(a.out);
  C.M = a.M;
  C.g = a.g;
  C.j0 = a.j0;
  C.n = a.n;
  C.Nm = a.n >> a.j0;
  C.baseD = inv_off(L, a.M, a.g + 1) * R;
  const int e = 2 * int(threadIdx.x);
  C.active = e < a.M * R;
  C.mm = e / R;
  C.r = e % R;
  const int64_t nrb = (int64_t(1) << a.j0) / R;
  const int64_t rb = int64_t(blockIdx.x) % nrb, seg = int64_t(blockIdx.x) / nrb;
  C.r0 = rb * R;
  C.mlo = seg * a.P;
  C.mhi = C.mlo + a.P;
  const int64_t cbot = C.mlo - a.warm;
#pragma unroll
  for (int l = 0; l <= kGmax; ++l) C.head[l] = 0;

  // prologue: chunk 0 of every ring
  C.template load_next<1>(cbot);
  C.template park<1>();
  __syncthreads();

  for (int k = 0; k < a.nchunks; ++k) {
    C.c = cbot + int64_t(k) * a.M;
    const bool have_next = (k + 1 < a.nchunks);
    if (have_next) C.template load_next<1>(C.c + a.M);
    C.template levels_down<kGmax>();
    // advance rings (ascending march: head += M mod Q)
#pragma unroll
    for (int l = 1; l <= kGmax; ++l) {
      const int Q = a.M + inv_win(L, l);
      int h = C.head[l] + a.M;
      while (h >= Q) h -= Q;
      C.head[l] = h;
    }
    if (have_next) {
      C.template park<1>();
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ host

namespace {

constexpr int kNT1 = 512;   // threads, group starting at level 1 (R = 1)
constexpr int kNTR = 256;   // threads, deeper groups (R = 32)
constexpr int kRdeep = 32;

struct Group {
  int j0, g, R, NT, M;
  int64_t P;
  int nseg, warm, nchunks;
  size_t lds;
};

size_t ring_elems(bool inverse, int L, int g, int M) {
  return inverse ? size_t(2) * inv_off(L, M, g + 1) : size_t(fwd_off(L, M, g));
}

int64_t round_up64(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Partition levels 1..J into groups.  The first group (R = 1) fuses as many
// levels as the LDS budget allows while its warm-up stays at most half a
// segment with one workgroup per CU; deeper groups run R = 32 residues.
std::vector<Group> plan_groups(bool inverse, int64_t n, int J, int L, size_t esz) {
  const DeviceInfo& dev = device_info();
  const size_t budget = std::min<size_t>(dev.max_lds_per_block, 160 * 1024) - 1024;
  const int64_t cus = std::max(1, dev.num_cus);
  std::vector<Group> gs;
  int j0 = 0;
  while (j0 < J) {
    Group G{};
    G.j0 = j0;
    const int64_t B = int64_t(1) << j0;
    // R = 1 for the first group; deeper groups take 32 residues when the
    // stride allows it (B >= 32), else all B of them
    G.R = (j0 == 0) ? 1 : int(std::min<int64_t>(B, kRdeep));
    G.NT = (G.R == 1) ? kNT1 : kNTR;
    const int64_t Nm = n >> j0;
    G.M = int(std::min<int64_t>(2 * G.NT / G.R, Nm));
    const int64_t nrb = B / G.R;
    const int64_t target = std::max<int64_t>(1, cus / nrb);
    int64_t nseg_t = 1;
    while (nseg_t * 2 <= target && Nm / (nseg_t * 2) >= G.M) nseg_t *= 2;
    const int64_t Pt = Nm / nseg_t;
    // the first group always fuses at least min(J, 5) levels so that deeper
    // groups start at a stride of 32 or more
    const int gmin = (j0 == 0) ? std::min(J, 5) : 1;
    int gbest = gmin;
    for (int g = gmin + 1; g <= std::min(J - j0, kGmax); ++g) {
      if (ring_elems(inverse, L, g, G.M) * G.R * esz > budget) break;
      int64_t warm = round_up64(int64_t(L - 1) * ((int64_t(1) << g) - 1), G.M);
      if (warm > Pt / 2 && nseg_t > 1) break;
      gbest = g;
    }
    G.g = gbest;
    G.warm = int(round_up64(int64_t(L - 1) * ((int64_t(1) << G.g) - 1), G.M));
    int64_t nseg = 1;
    while (nseg * 2 <= target && Nm / (nseg * 2) >= G.M &&
           Nm / (nseg * 2) >= 2 * int64_t(G.warm))
      nseg *= 2;
    G.nseg = int(nseg);
    G.P = Nm / nseg;
    G.nchunks = int((G.P + G.warm) / G.M);
    G.lds = ring_elems(inverse, L, G.g, G.M) * G.R * esz;
    gs.push_back(G);
    j0 += G.g;
  }
  return gs;
}

template <typename KERN>
void allow_big_lds(KERN kern) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <typename T, int FILT, int R, int NT, bool INV>
int launch_group(const Group& G, const ModwtArgs& a, hipStream_t st) {
  auto kern = INV ? modwt_inv_v2<T, FILT, R, NT> : modwt_fwd_v2<T, FILT, R, NT>;
  static std::once_flag once;
  std::call_once(once, [&] { allow_big_lds(kern); });
  int64_t nrb = (int64_t(1) << G.j0) / G.R;
  int64_t blocks = nrb * G.nseg;
  if (blocks > 0x7fffffff) return set_error(WH_ERR_ARG, "modwt: grid too large");
  hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(NT), G.lds, st, a);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

template <typename T, int FILT, bool INV>
int launch_r(const Group& G, const ModwtArgs& a, hipStream_t st) {
  switch (G.R) {
    case 1: return launch_group<T, FILT, 1, kNT1, INV>(G, a, st);
    case 2: return launch_group<T, FILT, 2, kNTR, INV>(G, a, st);
    case 4: return launch_group<T, FILT, 4, kNTR, INV>(G, a, st);
    case 8: return launch_group<T, FILT, 8, kNTR, INV>(G, a, st);
    case 16: return launch_group<T, FILT, 16, kNTR, INV>(G, a, st);
    default: return launch_group<T, FILT, kRdeep, kNTR, INV>(G, a, st);
  }
}

template <typename T, bool INV>
int launch_any(int filt, const Group& G, const ModwtArgs& a, hipStream_t st) {
  switch (filt) {
    case kHAAR: return launch_r<T, kHAAR, INV>(G, a, st);
    case kD4: return launch_r<T, kD4, INV>(G, a, st);
    case kC6: return launch_r<T, kC6, INV>(G, a, st);
    default: return launch_r<T, kLA8, INV>(G, a, st);
  }
}

ModwtArgs args_of(const Group& G, int64_t n) {
  ModwtArgs a{};
  a.n = n;
  a.P = G.P;
  a.j0 = G.j0;
  a.g = G.g;
  a.M = G.M;
  a.warm = G.warm;
  a.nchunks = G.nchunks;
  return a;
}

template <typename T>
int modwt_fwd_t(const void* x, void* xmod, int64_t n, int J, int filt, hipStream_t st) {
  const int L = filter_length(filt);
  auto groups = plan_groups(false, n, J, L, sizeof(T));
  for (const Group& G : groups) {
    ModwtArgs a = args_of(G, n);
    a.in = (G.j0 == 0) ? x : static_cast<const T*>(xmod) + 2 * n * int64_t(G.j0 - 1);
    a.xmod = xmod;
    int rc = launch_any<T, false>(filt, G, a, st);
    if (rc) return rc;
  }
  return WH_OK;
}

template <typename T>
int modwt_inv_t(const void* xmod, void* x, int64_t n, int J, int filt, const wh_thresh* thr,
                hipStream_t st) {
  const int L = filter_length(filt);
  auto groups = plan_groups(true, n, J, L, sizeof(T));
  // the S_{j0} handed between groups lives in library scratch (2 n values)
  T* scratch = nullptr;
  if (groups.size() > 1) {
    void* p = nullptr;
    int rc = scratch_get(2 * size_t(n) * sizeof(T), &p);
    if (rc) return rc;
    scratch = static_cast<T*>(p);
  }
  const T* stop = static_cast<const T*>(xmod) + 2 * n * int64_t(J - 1);
  for (int gi = int(groups.size()) - 1; gi >= 0; --gi) {
    const Group& G = groups[gi];
    ModwtArgs a = args_of(G, n);
    a.xmod = const_cast<void*>(xmod);
    a.stop = stop;
    T* dst = (G.j0 == 0) ? static_cast<T*>(x) : scratch + (gi & 1) * n;
    a.out = dst;
    if (thr && thr->enabled) {
      a.th_on = 1;
      a.hardness = thr->hardness;
      a.minlevel = thr->minlevel;
      a.maxlevel = thr->maxlevel;
      a.th_value = thr->value;
      a.th_levels = thr->levels;
    }
    int rc = launch_any<T, true>(filt, G, a, st);
    if (rc) return rc;
    stop = dst;
  }
  return WH_OK;
}

}  // namespace

int modwt_forward(const void* x, void* xmod, int64_t n, int J, int filt, int dtype,
                  hipStream_t st) {
  return dtype == WH_F32 ? modwt_fwd_t<float>(x, xmod, n, J, filt, st)
                         : modwt_fwd_t<double>(x, xmod, n, J, filt, st);
}

int modwt_inverse(const void* xmod, void* x, int64_t n, int J, int filt, int dtype,
                  const wh_thresh* thr, hipStream_t st) {
  return dtype == WH_F32 ? modwt_inv_t<float>(xmod, x, n, J, filt, thr, st)
                         : modwt_inv_t<double>(xmod, x, n, J, filt, thr, st);
}

std::string modwt_plan_string(bool inverse, int64_t n, int J, int filt, int dtype) {
  const int L = filter_length(filt);
  size_t esz = dtype == WH_F32 ? 4 : 8;
  auto gs = plan_groups(inverse, n, J, L, esz);
  std::string s;
  for (auto& G : gs) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "{j0=%d g=%d R=%d NT=%d M=%d nseg=%d P=%lld warm=%d chunks=%d lds=%zu} ", G.j0,
             G.g, G.R, G.NT, G.M, G.nseg, (long long)G.P, G.warm, G.nchunks, G.lds);
    s += buf;
  }
  return s;
}

}  // namespace wh
