// modwt_kern.h -- device side of the level-fused time-ordered MODWT.
//
// Math (reference src/daub4.cpp:561-589, src/haar.cpp:590-616; C6/LA8
// use the same pattern): with s = 2^{j-1}
//   S_j[i] = sum_k C_k S_{j-1}[(i + k s) mod n],  D_j[i] = sum_k g_k S_{j-1}[...]
//   S_{j-1}[i] = 1/2 sum_k (C_k S_j[(i - k s) mod n] + g_k D_j[(i - k s) mod n])
// stored as xmod[2n(j-1) + i] = S_j[i], xmod[2n(j-1) + n + i] = D_j[i].
//
// Kernel design.  A "group" fuses G consecutive levels j0+1..j0+G.  Level
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
// (L-1)(2^G - 1) positions (rounded up to M) whose outputs are computed but
// not stored -- the only redundant work.
//
// Each thread owns two elements (four in the first inverse group's levels,
// see InvLinCtx).  For R = 1 they are the outputs q and q + s of one
// residue class of the level's stride s, whose L-tap windows overlap, so a
// level reads L + 1 values per thread instead of 2L (for s = 1 the two are
// adjacent and read as 16-byte pairs); for R > 1 they are two adjacent
// residues and every access is a 16-byte vector.  Ring slots
// are kept even so a pair never straddles the ring's wrap; the inverse
// interleaves S and D per slot (R = 1) so one 16-byte read returns both.
//
// Latency hiding: the next chunk's HBM loads are issued before a chunk's
// arithmetic and parked in LDS after it.  G is a template parameter and the
// warm-up chunks run a loop of their own, so every load and store of a
// chunk is unconditional: the compiler can then count the stores issued
// after a prefetch (CDNA's vmcnt counts loads and stores together) instead
// of draining the whole memory pipeline before parking it.
#pragma once

#include "wh_common.h"

namespace wh {

struct ModwtArgs {
  const void* in;        // fwd: group input S_{j0} (x when j0 == 0)
  void* xmod;            // fwd: output blocks; inv: detail blocks
  const void* stop;      // inv: S_{j0+G} (xmod slot or scratch)
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
  // signal batches: blockIdx.y selects the signal, offset by these strides
  // (elements) from in / xmod / stop / out
  int64_t bs_in, bs_xmod, bs_stop, bs_out;
};

// the arguments of the signal handled by this workgroup (blockIdx.y)
template <typename T>
__device__ __forceinline__ ModwtArgs batch_slice(const ModwtArgs& a) {
  ModwtArgs b = a;
  const int64_t y = blockIdx.y;
  b.in = static_cast<const T*>(a.in) + y * a.bs_in;
  b.xmod = static_cast<T*>(a.xmod) + y * a.bs_xmod;
  b.stop = static_cast<const T*>(a.stop) + y * a.bs_stop;
  b.out = static_cast<T*>(a.out) + y * a.bs_out;
  return b;
}

constexpr int kGmax = 10;

__host__ __device__ constexpr int even_up(int v) { return (v + 1) & ~1; }

// Ring geometry in m units.  Forward ring l (l = 0..G-1) holds the input of
// level l+1, window (L-1) 2^l; inverse ring l (l = 1..G) holds S_l and D_l,
// window (L-1) 2^{l-1}.  Windows are rounded up to even.
__host__ __device__ constexpr int fwd_win(int L, int l) { return even_up((L - 1) << l); }
__host__ __device__ constexpr int inv_win(int L, int l) { return even_up((L - 1) << (l - 1)); }
__host__ __device__ constexpr int fwd_win_sum(int L, int l) {
  return l == 0 ? 0 : fwd_win_sum(L, l - 1) + fwd_win(L, l - 1);
}
__host__ __device__ constexpr int inv_win_sum(int L, int l) {  // rings 1..l-1
  return l <= 1 ? 0 : inv_win_sum(L, l - 1) + inv_win(L, l - 1);
}
__host__ __device__ inline int fwd_off(int L, int M, int l) { return l * M + fwd_win_sum(L, l); }
__host__ __device__ inline int inv_off(int L, int M, int l) {
  return (l - 1) * M + inv_win_sum(L, l);
}

// Linear buffers of the first inverse group (InvLinCtx).  Level l holds the
// positions q in [-W_l, M) of the chunk, p = q + W_l.  Levels of stride
// s = 2^{l-1} <= 8 with M >= 4 s are "planar": position p lives in plane
// 4 (p mod s) + (p / s mod 4) at index p / (4 s), so the four-outputs-per-
// thread mapping reads consecutive slots across lanes (no LDS bank
// conflicts).  Other levels keep position order (slot p).
__host__ __device__ constexpr bool lin_planar(int M, int l) {
  return l <= 4 && M >= (4 << (l - 1));
}
__host__ __device__ inline int lin_plane(int L, int M, int l) {  // slots per plane
  return ((inv_win(L, l) + M) / (1 << (l - 1)) + 3) / 4;
}
__host__ __device__ inline int lin_size(int L, int M, int l) {
  return lin_planar(M, l) ? (4 << (l - 1)) * lin_plane(L, M, l) : inv_win(L, l) + M;
}
__host__ __device__ inline int lin_off(int L, int M, int l) {  // levels 1..l-1
  int o = 0;
  for (int i = 1; i < l; ++i) o += lin_size(L, M, i);
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
// streaming read: the inverse reads every coefficient once, so its loads
// should not displace other data (such as the signal) from the caches
// (one vector access per pair: p is aligned to two elements)
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type ldv_nt(const T* p) {
  typedef T ev2 __attribute__((ext_vector_type(2)));
  const ev2 r = __builtin_nontemporal_load(reinterpret_cast<const ev2*>(p));
  return {r.x, r.y};
}
template <typename T>
__device__ __forceinline__ void stv_nt(T* p, typename Vec2<T>::type v) {
  typedef T ev2 __attribute__((ext_vector_type(2)));
  const ev2 r = {v.x, v.y};
  __builtin_nontemporal_store(r, reinterpret_cast<ev2*>(p));
}

// analysis outputs of taps w[0..L-1] and w[1..L] (reference term order)
template <typename T, int FILT>
__device__ __forceinline__ void analysis2(const T* w, T& s0, T& d0, T& s1, T& d1) {
  constexpr int L = Filter<FILT>::L;
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
}

// 1/2 sum_k C_k S[tap k] + g_k D[tap k] in the reference's term order
// k = L-2, L-4, ..., 0, L-1, ..., 1 (src/daub4.cpp:578-583); tap k of the
// output is (ws[base - k], wd[base - k])
template <typename T, int FILT>
__device__ __forceinline__ T synthesis(const T* ws, const T* wd, int base) {
  constexpr int L = Filter<FILT>::L;
  if constexpr (FILT == kHAAR) {
    return (ws[base] + wd[base] + ws[base - 1] - wd[base - 1]) * T(kHInvR2);
  } else {
    T e = ws[base - (L - 2)] * T(Filter<FILT>::c(L - 2));
    e = fma(wd[base - (L - 2)], T(gcoef<FILT>(L - 2)), e);
#pragma unroll
    for (int kk = L - 4; kk >= 0; kk -= 2) {
      e = fma(ws[base - kk], T(Filter<FILT>::c(kk)), e);
      e = fma(wd[base - kk], T(gcoef<FILT>(kk)), e);
    }
#pragma unroll
    for (int kk = L - 1; kk >= 1; kk -= 2) {
      e = fma(ws[base - kk], T(Filter<FILT>::c(kk)), e);
      e = fma(wd[base - kk], T(gcoef<FILT>(kk)), e);
    }
    return T(0.5) * e;
  }
}

// ---------------------------------------------------------------- forward
// Store policy of the forward: kStoreAll writes S_j and D_j of every level
// with streaming stores (the transform's output); kStoreSmooth writes only
// D_j and the last S with ordinary stores (a smoothing pipeline's
// intermediate, which its inverse re-reads while it may still sit in L2);
// kStoreAdjoint makes the forward the adjoint of the inverse: level l of the
// group stores 2^-l D and, at the group's last level, 2^-G S, and zeros
// in the other S slots (the inverse never reads them).  The scalings are
// powers of two, hence exact.  kStoreSmoothHist stores as kStoreSmooth and
// also counts the first radix digit (sel_digit0) of every D_1 value into
// an LDS histogram, when the body is given one (the batched pipeline's
// median of D_1 then starts from it).
constexpr int kStoreAll = 0;
constexpr int kStoreSmooth = 1;
constexpr int kStoreAdjoint = 2;
constexpr int kStoreSmoothHist = 3;

template <typename T>
__device__ __forceinline__ void st_pol(T* p, T v, bool nt) {
  if (nt)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

template <typename T, int FILT, int R, int G, int MAP, int SP = kStoreAll>
struct FwdCtx {
  using V = typename Vec2<T>::type;
  static constexpr int L = Filter<FILT>::L;
  T* lds;
  T* xmod;
  int M, mm, r, j0;
  int64_t n, r0, c;
  int head[G];  // ring heads (slot of position c), ring l = 0..G-1
  unsigned* hist1;  // kStoreSmoothHist: digit counts of D_1, or null
  bool own;         // this thread's outputs are not another thread's duplicates

  // level l (1-based within the group): ring l-1 -> S_{j0+l}, D_{j0+l}
  template <int l, bool STORE>
  __device__ __forceinline__ void level() {
    constexpr int s = 1 << (l - 1);
    const int Qin = M + fwd_win(L, l - 1);
    const T* rin = lds + size_t(fwd_off(L, M, l - 1)) * R;
    T s0, d0, s1, d1;
    if constexpr (R == 1 && (MAP == 1 || s == 1)) {
      // outputs q0 and q0 + s of one residue class; taps q0 + u s, u = 0..L
      const int t = mm >> 1;
      const int q0 = (s == 1) ? mm : 2 * s * (t / s) + (t % s);
      T w[L + 2];
      int sl = head[l - 1] + q0;
      if (sl >= Qin) sl -= Qin;
      if constexpr (s == 1) {
#pragma unroll
        for (int q = 0; q <= L / 2; ++q) {
          int sq = sl + 2 * q;
          if (sq >= Qin) sq -= Qin;
          V v = ldv(rin + sq);
          w[2 * q] = v.x;
          w[2 * q + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int u = 0; u <= L; ++u) {
          int su = sl + u * s;
          if (su >= Qin) su -= Qin;
          w[u] = rin[su];
        }
      }
      analysis2<T, FILT>(w, s0, d0, s1, d1);
      if constexpr (l < G) {
        const int Qo = M + fwd_win(L, l);
        T* ro = lds + size_t(fwd_off(L, M, l));
        int so = head[l] + q0;
        if (so >= Qo) so -= Qo;
        if constexpr (s == 1) {
          stv(ro + so, V{s0, s1});
        } else {
          ro[so] = s0;
          so += s;
          if (so >= Qo) so -= Qo;
          ro[so] = s1;
        }
      }
      if constexpr (STORE && SP == kStoreAdjoint) {
        constexpr T sc = T(1.0 / double(1 << l));
        const T ks = (l == G) ? sc : T(0);
        T* oS = xmod + 2 * n * int64_t(j0 + l - 1) + (c + q0);
        if constexpr (s == 1) {
          stv_nt(oS, V{s0 * ks, s1 * ks});
          stv_nt(oS + n, V{d0 * sc, d1 * sc});
        } else {
          __builtin_nontemporal_store(s0 * ks, oS);
          __builtin_nontemporal_store(s1 * ks, oS + s);
          __builtin_nontemporal_store(d0 * sc, oS + n);
          __builtin_nontemporal_store(d1 * sc, oS + n + s);
        }
      } else if constexpr (STORE) {
        constexpr bool kS = (SP == kStoreAll) || (l == G);
        constexpr bool kNT = (SP == kStoreAll);
        T* oS = xmod + 2 * n * int64_t(j0 + l - 1) + (c + q0);
        if constexpr (s == 1) {
          if constexpr (kNT) {
            stv_nt(oS, V{s0, s1});
            stv_nt(oS + n, V{d0, d1});
          } else {
            if constexpr (kS) stv(oS, V{s0, s1});
            stv(oS + n, V{d0, d1});
          }
          if constexpr (SP == kStoreSmoothHist && l == 1) {
            if (hist1 && own && j0 == 0) {
              atomicAdd(&hist1[sel_digit0(d0)], 1u);
              atomicAdd(&hist1[sel_digit0(d1)], 1u);
            }
          }
        } else {
          if constexpr (kS) {
            st_pol(oS, s0, kNT);
            st_pol(oS + s, s1, kNT);
          }
          st_pol(oS + n, d0, kNT);
          st_pol(oS + n + s, d1, kNT);
        }
      }
    } else {
      // the pair is two adjacent elements (consecutive m for R = 1,
      // consecutive residues otherwise) read as 16-byte vectors
      T wx[L + 1], wy[L + 1];
      int sl = head[l - 1] + mm;
      if (sl >= Qin) sl -= Qin;
#pragma unroll
      for (int k = 0; k < L; ++k) {
        int sk = sl + k * s;
        if (sk >= Qin) sk -= Qin;
        V v = ldv(rin + size_t(sk) * R + r);
        wx[k] = v.x;
        wy[k] = v.y;
      }
      wx[L] = wx[L - 1];
      wy[L] = wy[L - 1];
      T u0, u1;
      analysis2<T, FILT>(wx, s0, d0, u0, u1);
      analysis2<T, FILT>(wy, s1, d1, u0, u1);
      if constexpr (l < G) {
        const int Qo = M + fwd_win(L, l);
        int so = head[l] + mm;
        if (so >= Qo) so -= Qo;
        stv(lds + size_t(fwd_off(L, M, l)) * R + size_t(so) * R + r, V{s0, s1});
      }
      if constexpr (STORE) {
        const int64_t i = (R == 1) ? (c + mm) : r0 + r + ((c + mm) << j0);
        T* oS = xmod + 2 * n * int64_t(j0 + l - 1) + i;
        if constexpr (SP == kStoreAll) {
          stv_nt(oS, V{s0, s1});
          stv_nt(oS + n, V{d0, d1});
        } else if constexpr (SP == kStoreAdjoint) {
          constexpr T sc = T(1.0 / double(1 << l));
          const T ks = (l == G) ? sc : T(0);
          stv_nt(oS, V{s0 * ks, s1 * ks});
          stv_nt(oS + n, V{d0 * sc, d1 * sc});
        } else {
          if constexpr (l == G) stv(oS, V{s0, s1});
          stv(oS + n, V{d0, d1});
        }
      }
    }
    __syncthreads();
  }

  template <int l, bool STORE>
  __device__ __forceinline__ void levels() {
    if constexpr (l <= G) {
      level<l, STORE>();
      levels<l + 1, STORE>();
    }
  }
  __device__ __forceinline__ void advance() {
#pragma unroll
    for (int l = 0; l < G; ++l) {
      const int Q = M + fwd_win(L, l);
      int h = head[l] - M;
      if (h < 0) h += Q;
      head[l] = h;
    }
  }
};

// One (residue block rb, segment seg) of a forward group, run by the whole
// workgroup of NT threads; smem holds the rings.
template <typename T, int FILT, int R, int NT, int G, int MAP, int SP = kStoreAll>
__device__ __forceinline__ void modwt_fwd_body(const ModwtArgs& a, int64_t rb, int64_t seg,
                                               unsigned char* smem_raw,
                                               unsigned* hist1 = nullptr) {
  using V = typename Vec2<T>::type;
  FwdCtx<T, FILT, R, G, MAP, SP> C;
  C.lds = reinterpret_cast<T*>(smem_raw);
  C.xmod = static_cast<T*>(a.xmod);
  C.M = a.M;
  C.j0 = a.j0;
  C.n = a.n;
  C.hist1 = hist1;
  // threads beyond the chunk (small signals) duplicate the work of thread
  // e mod M R and write identical values to identical places
  C.own = 2 * int(threadIdx.x) < a.M * R;
  const int e = (2 * int(threadIdx.x)) % (a.M * R);
  C.mm = e / R;
  C.r = e % R;
  const int64_t Nm = a.n >> a.j0;
  C.r0 = rb * R;
  const int64_t mhi = (seg + 1) * a.P;
  const T* __restrict__ in = static_cast<const T*>(a.in);
  const int64_t ctop = mhi + a.warm - a.M;
#pragma unroll
  for (int l = 0; l < G; ++l) C.head[l] = 0;

  auto in_ptr = [&](int64_t m) -> const T* {
    m &= (Nm - 1);
    return in + ((R == 1) ? m : C.r0 + C.r + (m << a.j0));
  };
  const int Q0 = a.M + fwd_win(Filter<FILT>::L, 0);
  // prologue: chunk 0 into ring 0
  stv(C.lds + size_t(C.mm) * R + C.r, ldv(in_ptr(ctop + C.mm)));
  __syncthreads();

  const int nwarm = a.warm / a.M;
  int k = 0;
  for (; k < nwarm; ++k) {
    C.c = ctop - int64_t(k) * a.M;
    const V pre = ldv(in_ptr(C.c - a.M + C.mm));
    C.template levels<1, false>();
    C.advance();
    int so = C.head[0] + C.mm;
    if (so >= Q0) so -= Q0;
    stv(C.lds + size_t(so) * R + C.r, pre);
    __syncthreads();
  }
  for (; k < a.nchunks; ++k) {
    C.c = ctop - int64_t(k) * a.M;
    const V pre = ldv(in_ptr(C.c - a.M + C.mm));  // wraps harmlessly on the last chunk
    C.template levels<1, true>();
    C.advance();
    int so = C.head[0] + C.mm;
    if (so >= Q0) so -= Q0;
    stv(C.lds + size_t(so) * R + C.r, pre);
    __syncthreads();
  }
}

template <typename T, int FILT, int R, int NT, int G, int MAP, int SP>
__global__ __launch_bounds__(NT) void modwt_fwd_k(ModwtArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t nrb = (int64_t(1) << a.j0) / R;
  const ModwtArgs b = batch_slice<T>(a);
  modwt_fwd_body<T, FILT, R, NT, G, MAP, SP>(b, int64_t(blockIdx.x) % nrb,
                                             int64_t(blockIdx.x) / nrb, smem_raw);
}

// ---------------------------------------------------------------- inverse
// Ring l (l = 1..G) holds S_l and D_l: interleaved per slot (S, D) for the
// residue-class mapping, otherwise the S rings come first and the D rings
// follow at baseD.
template <typename T, int FILT, int R, int G, int MAP>
struct InvCtx {
  static_assert(R > 1 || MAP == 1, "the R = 1 inverse uses the residue-class mapping");
  using V = typename Vec2<T>::type;
  static constexpr int L = Filter<FILT>::L;
  T* lds;
  int baseD;
  const ModwtArgs* a;
  const T* xmod;
  const T* stop;
  T* out;
  int M, mm, r, j0;
  int64_t n, Nm, r0, c;
  int head[G + 1];  // head of ring l (1..G): slot of position c - W'_l
  V pre[G + 1];     // next chunk: D_l (l = 1..G) and S_G (index 0)
  T thv[G + 1];     // threshold of level l, negative where none applies

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
  // (S, D) interleaved per slot only for the residue-class mapping, whose
  // lanes read one position each; the adjacent-pair layouts keep S and D
  // apart so a 16-byte read returns two consecutive slots of one array
  static constexpr bool kInter = (R == 1 && MAP == 1);
  __device__ __forceinline__ int sidx(int l, int sl) const {
    return kInter ? 2 * (inv_off(L, M, l) + sl) : (inv_off(L, M, l) + sl) * R + r;
  }
  __device__ __forceinline__ int didx(int l, int sl) const {
    return kInter ? 2 * (inv_off(L, M, l) + sl) + 1
                  : baseD + (inv_off(L, M, l) + sl) * R + r;
  }

  // HBM loads of chunk cc for every ring
  __device__ __forceinline__ void load_next(int64_t cc) {
    const int64_t i = gidx(cc + mm);
#pragma unroll
    for (int l = 1; l <= G; ++l) pre[l] = ldv_nt(xmod + 2 * n * int64_t(j0 + l - 1) + n + i);
    pre[0] = ldv_nt(stop + i);
  }
  // park the loaded chunk in the rings (slots head_l + W'_l + mm)
  __device__ __forceinline__ void park() {
#pragma unroll
    for (int l = 1; l <= G; ++l) {
      const int Q = M + inv_win(L, l);
      int so = head[l] + inv_win(L, l) + mm;
      if (so >= Q) so -= Q;
      const V d = thr2(pre[l], thv[l]);
      if constexpr (kInter) {
        // slots so, so+1 (so even, Q even: no wrap between them)
        T* p = lds + sidx(l, so);
        if (l == G) {
          stv(p, V{pre[0].x, d.x});
          stv(p + 2, V{pre[0].y, d.y});
        } else {
          p[1] = d.x;
          p[3] = d.y;
        }
      } else {
        stv(lds + didx(l, so), d);
        if (l == G) stv(lds + sidx(l, so), pre[0]);
      }
    }
  }

  // level l: S_l, D_l (ring l) -> S_{l-1} (ring l-1, or HBM when l == 1)
  template <int l, bool STORE>
  __device__ __forceinline__ void level() {
    constexpr int s = 1 << (l - 1);
    const int W = inv_win(L, l);
    const int Q = M + W;
    T e0, e1;
    int q0;  // offset of the first output from c; the second is q0 + s
    if constexpr (R == 1 && (MAP == 1 || s == 1)) {
      const int t = mm >> 1;
      q0 = (s == 1) ? mm : 2 * s * (t / s) + (t % s);
      // taps of outputs q0 and q0 + s are positions q0 + u s, u = -(L-1)..1,
      // kept at index u + L - 1; position p sits at slot head + W + (p - c)
      T ws[L + 1], wd[L + 1];
      int sl = (s == 1) ? head[l] + q0 + 1 : head[l] + W + q0 - (L - 1) * s;
      if (sl >= Q) sl -= Q;
#pragma unroll
      for (int u = 0; u <= L; ++u) {
        int su = sl + u * s;
        if (su >= Q) su -= Q;
        V v = ldv(lds + sidx(l, su));
        ws[u] = v.x;
        wd[u] = v.y;
      }
      e0 = synthesis<T, FILT>(ws, wd, L - 1);
      e1 = synthesis<T, FILT>(ws, wd, L);
    } else {
      q0 = mm;
      int sl = head[l] + W + mm;
      if (sl >= Q) sl -= Q;
      T ws0[L], wd0[L], ws1[L], wd1[L];
#pragma unroll
      for (int kk = 0; kk < L; ++kk) {
        int sk = sl - kk * s;
        if (sk < 0) sk += Q;
        V vs = ldv(lds + sidx(l, sk)), vd = ldv(lds + didx(l, sk));
        ws0[L - 1 - kk] = vs.x;
        ws1[L - 1 - kk] = vs.y;
        wd0[L - 1 - kk] = vd.x;
        wd1[L - 1 - kk] = vd.y;
      }
      e0 = synthesis<T, FILT>(ws0, wd0, L - 1);
      e1 = synthesis<T, FILT>(ws1, wd1, L - 1);
    }
    if constexpr (l > 1) {
      const int Wo = inv_win(L, l - 1);
      const int Qo = M + Wo;
      int so = head[l - 1] + Wo + q0;
      if (so >= Qo) so -= Qo;
      if constexpr (R == 1 && MAP == 1) {
        lds[sidx(l - 1, so)] = e0;
        so += s;
        if (so >= Qo) so -= Qo;
        lds[sidx(l - 1, so)] = e1;
      } else {
        stv(lds + sidx(l - 1, so), V{e0, e1});
      }
    } else if constexpr (STORE) {
      const int64_t m = c + q0;
      const int64_t i = (R == 1) ? m : r0 + r + (m << j0);
      // a deeper group (R > 1) hands S_{j0} to the next group launched,
      // which reads it right away: ordinary stores leave it in the
      // last-level cache (measured: the inverse 1 % faster at LA8, J = 10,
      // n = 2^24); the transform's own output streams past the caches
      if constexpr (R > 1)
        stv(out + i, V{e0, e1});
      else
        stv_nt(out + i, V{e0, e1});
    }
    __syncthreads();
  }

  template <int l, bool STORE>
  __device__ __forceinline__ void levels() {
    if constexpr (l >= 1) {
      level<l, STORE>();
      levels<l - 1, STORE>();
    }
  }
  __device__ __forceinline__ void advance() {
#pragma unroll
    for (int l = 1; l <= G; ++l) {
      const int Q = M + inv_win(L, l);
      int h = head[l] + M;
      if (h >= Q) h -= Q;
      head[l] = h;
    }
  }
};

template <typename T, int FILT, int R, int NT, int G, int MAP>
__device__ __forceinline__ void modwt_inv_body(const ModwtArgs& a, int64_t rb, int64_t seg,
                                               unsigned char* smem_raw) {
  constexpr int L = Filter<FILT>::L;
  InvCtx<T, FILT, R, G, MAP> C;
  C.lds = reinterpret_cast<T*>(smem_raw);
  C.a = &a;
  C.xmod = static_cast<const T*>(a.xmod);
  C.stop = static_cast<const T*>(a.stop);
  C.out = static_cast<T*>(a.out);
  C.M = a.M;
  C.j0 = a.j0;
  C.n = a.n;
  C.Nm = a.n >> a.j0;
  C.baseD = (R == 1 && MAP == 1) ? 0 : inv_off(L, a.M, G + 1) * R;
  const int e = (2 * int(threadIdx.x)) % (a.M * R);
  C.mm = e / R;
  C.r = e % R;
  C.r0 = rb * R;
  const int64_t cbot = seg * a.P - a.warm;
#pragma unroll
  for (int l = 0; l <= G; ++l) C.head[l] = 0;
#pragma unroll
  for (int l = 1; l <= G; ++l) C.thv[l] = C.thr(l);  // read once, outside the chunk loop

  // prologue: chunk 0 of every ring
  C.load_next(cbot);
  C.park();
  __syncthreads();

  const int nwarm = a.warm / a.M;
  int k = 0;
  for (; k < nwarm; ++k) {
    C.c = cbot + int64_t(k) * a.M;
    C.load_next(C.c + a.M);
    C.template levels<G, false>();
    C.advance();
    C.park();
    __syncthreads();
  }
  for (; k < a.nchunks; ++k) {
    C.c = cbot + int64_t(k) * a.M;
    C.load_next(C.c + a.M);  // wraps harmlessly on the last chunk
    C.template levels<G, true>();
    C.advance();
    C.park();
    __syncthreads();
  }
}

// ------------------------------------------------- inverse, linear buffers
// The R = 1 group with the residue-class mapping, but each level keeps a
// linear buffer instead of a ring: it holds the W_l positions before the
// chunk (the window) and the chunk's M positions, so every tap of a thread
// is a fixed offset from a per-thread base address -- no wrap arithmetic
// per tap.  At the end of a chunk the last W_l positions move to the window
// (through registers, since the regions overlap when W_l > M) before the
// next chunk is parked behind them.  Slots interleave (S, D).  Levels of
// stride up to 8 use the planar slot order of lin_planar (conflict-free
// reads for the four-output mapping); deeper levels keep position order,
// where lanes of one residue class already read consecutive slots.
template <typename T, int FILT, int G, int NT>
struct InvLinCtx {
  using V = typename Vec2<T>::type;
  static constexpr int L = Filter<FILT>::L;
  T* lds;
  const ModwtArgs* a;
  const T* xmod;
  const T* stop;
  T* out;
  int M, mm, lgM;
  int64_t n, c;
  int64_t cst;      // where this chunk's outputs go (c; a warm-up chunk stores to the
                    // segment's first chunk, whose own outputs later replace them)
  int offs[G + 1];  // first slot of level l's buffer
  int ps[5];        // slots per plane of the planar levels 1..4, 0 if not planar
  V pre[G + 1];     // next chunk: D_l (l = 1..G) and S_G (index 0)
  T thv[G + 1];     // threshold of level l, negative where none applies

  __device__ __forceinline__ void init_layout() {
    lgM = 31 - __clz(M);
    int o = 0;
#pragma unroll
    for (int l = 1; l <= G; ++l) {
      offs[l] = o;
      o += lin_size(L, M, l);
    }
#pragma unroll
    for (int l = 1; l <= 4; ++l) ps[l] = lin_planar(M, l) ? lin_plane(L, M, l) : 0;
#pragma unroll
    for (int l = 1; l <= G; ++l) thv[l] = thr(l);  // read once, outside the chunk loop
  }
  template <int l>
  __device__ __forceinline__ bool planar() const {
    if constexpr (l <= 4) return ps[l] != 0;
    return false;
  }
  // slot of chunk position q in [-W_l, M) of level l
  template <int l>
  __device__ __forceinline__ int slot(int q) const {
    const int p = q + inv_win(L, l);
    if constexpr (l <= 4) {
      if (ps[l]) {
        const int r = p & ((1 << (l - 1)) - 1), cc = p >> (l - 1);
        return offs[l] + ((r << 2) | (cc & 3)) * ps[l] + (cc >> 2);
      }
    }
    return offs[l] + p;
  }
  __device__ __forceinline__ T thr(int l) const {
    const int lev = l - 1;
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
  __device__ __forceinline__ void load_next(int64_t cc) {
    const int64_t i = (cc + mm) & (n - 1);
#pragma unroll
    for (int l = 1; l <= G; ++l) pre[l] = ldv_nt(xmod + 2 * n * int64_t(l - 1) + n + i);
    pre[0] = ldv_nt(stop + i);
  }
  // the loaded chunk goes to positions mm, mm + 1 of every level
  template <int l>
  __device__ __forceinline__ void park_level() {
    const V d = thr2(pre[l], thv[l]);
    if (planar<l>()) {
      T* p0 = lds + 2 * slot<l>(mm);
      T* p1 = lds + 2 * slot<l>(mm + 1);
      if constexpr (l == G) {
        stv(p0, V{pre[0].x, d.x});
        stv(p1, V{pre[0].y, d.y});
      } else {
        p0[1] = d.x;
        p1[1] = d.y;
      }
    } else {
      T* p = lds + 2 * slot<l>(mm);
      if constexpr (l == G) {
        stv(p, V{pre[0].x, d.x});
        stv(p + 2, V{pre[0].y, d.y});
      } else {
        p[1] = d.x;
        p[3] = d.y;
      }
    }
    if constexpr (l < G) park_level<l + 1>();
  }
  __device__ __forceinline__ void park() { park_level<1>(); }

  // S_{l-1} at chunk position q (level l > 1) or the output x (level 1)
  template <int l>
  __device__ __forceinline__ void put(int q, T e) {
    lds[2 * slot<l - 1>(q)] = e;
  }

  // Planar level: each of M/4 threads takes four outputs r + s (4j + i),
  // i = 0..3, of residue class r; their taps r + s (4j + u + e), u = 0..L+2,
  // are L + 3 slot reads for four outputs.  Lanes of one class read
  // consecutive slots of one plane.
  template <int l>
  __device__ __forceinline__ void level4p() {
    constexpr int s = 1 << (l - 1);
    constexpr int e = inv_win(L, l) / s - (L - 1);  // 1 where the window was rounded up
    const int tt = threadIdx.x;
    if (tt < (M >> 2)) {
      const int sh = lgM - l - 1;  // log2 of the threads per class, M / (4 s)
      const int r = tt >> sh, j = tt & ((1 << sh) - 1);
      const T* b = lds + 2 * (offs[l] + (r << 2) * ps[l] + j);
      T ws[L + 3], wd[L + 3];
#pragma unroll
      for (int u = 0; u < L + 3; ++u) {
        const int cu = u + e;
        const V v = ldv(b + 2 * ((cu & 3) * ps[l] + (cu >> 2)));
        ws[u] = v.x;
        wd[u] = v.y;
      }
      T o[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) o[i] = synthesis<T, FILT>(ws, wd, L - 1 + i);
      if constexpr (l > 1) {
        const int q0 = r + 4 * s * j;
#pragma unroll
        for (int i = 0; i < 4; ++i) put<l>(q0 + i * s, o[i]);
      } else {
        stv_nt(out + cst + 4 * j, V{o[0], o[1]});
        stv_nt(out + cst + 4 * j + 2, V{o[2], o[3]});
      }
    }
    __syncthreads();
  }

  // Position-ordered level of stride 16 or more: the same four-output
  // mapping with lanes over consecutive residues; needs M a multiple of 4 s.
  template <int l>
  __device__ __forceinline__ void level4() {
    constexpr int s = 1 << (l - 1);
    const int tt = threadIdx.x;
    if (tt < (M >> 2)) {
      const int q0 = 4 * s * (tt / s) + (tt % s);
      const T* b = lds + 2 * slot<l>(q0 - (L - 1) * s);
      T ws[L + 3], wd[L + 3];
#pragma unroll
      for (int u = 0; u < L + 3; ++u) {
        const V v = ldv(b + 2 * u * s);
        ws[u] = v.x;
        wd[u] = v.y;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) put<l>(q0 + j * s, synthesis<T, FILT>(ws, wd, L - 1 + j));
    }
    __syncthreads();
  }

  template <int l>
  __device__ __forceinline__ void level() {
    constexpr int s = 1 << (l - 1);
    if constexpr (l <= 4) {
      if (planar<l>()) {
        level4p<l>();
        return;
      }
    }
    if constexpr (s >= 16) {
      if ((M & (4 * s - 1)) == 0) {
        level4<l>();
        return;
      }
    }
    // small chunks: two outputs q0, q0 + s of one residue class per thread,
    // taps q0 + u s, u = -(L-1)..1
    const int t = mm >> 1;
    const int q0 = (s == 1) ? mm : 2 * s * (t / s) + (t % s);
    const T* b = lds + 2 * slot<l>(q0 - (L - 1) * s);
    T ws[L + 1], wd[L + 1];
#pragma unroll
    for (int u = 0; u <= L; ++u) {
      const V v = ldv(b + 2 * u * s);
      ws[u] = v.x;
      wd[u] = v.y;
    }
    const T e0 = synthesis<T, FILT>(ws, wd, L - 1);
    const T e1 = synthesis<T, FILT>(ws, wd, L);
    if constexpr (l > 1) {
      put<l>(q0, e0);
      put<l>(q0 + s, e1);
    } else {
      stv_nt(out + cst + q0, V{e0, e1});
    }
    __syncthreads();
  }
  template <int l>
  __device__ __forceinline__ void levels() {
    if constexpr (l >= 1) {
      level<l>();
      levels<l - 1>();
    }
  }
  // move positions [M - W_l, M) to [-W_l, 0) for every level, then park
  // the registers of buffer B
  template <int l>
  __device__ __forceinline__ void shift_and_park() {
    if constexpr (l > G) {
      __syncthreads();  // every old slot read before any is overwritten
      park();
    } else {
      constexpr int W = inv_win(L, l);
      constexpr int K = (W + NT - 1) / NT;
      // a position and the one M later share a plane (M is a multiple of
      // 4 s) or are M slots apart
      const int dist = planar<l>() ? (M >> (l + 1)) : M;
      V tmp[K];
      int dst[K];
      const int tid = threadIdx.x;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = tid + k * NT;
        dst[k] = slot<l>(i - W);
        if (i < W) tmp[k] = ldv(lds + 2 * (dst[k] + dist));
      }
      shift_and_park<l + 1>();
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int i = tid + k * NT;
        if (i < W) stv(lds + 2 * dst[k], tmp[k]);
      }
    }
  }
};

template <typename T, int FILT, int NT, int G>
__device__ __forceinline__ void modwt_inv_lin_body(const ModwtArgs& a, int64_t seg,
                                                   unsigned char* smem_raw) {
  InvLinCtx<T, FILT, G, NT> C;
  C.lds = reinterpret_cast<T*>(smem_raw);
  C.a = &a;
  C.xmod = static_cast<const T*>(a.xmod);
  C.stop = static_cast<const T*>(a.stop);
  C.out = static_cast<T*>(a.out);
  C.M = a.M;
  C.n = a.n;
  C.mm = (2 * int(threadIdx.x)) % a.M;
  C.init_layout();
  const int64_t cbot = seg * a.P - a.warm;

  // chunk k + 1's coefficients are loaded while chunk k is computed (the
  // positions wrap harmlessly past the last chunk).  Warm-up chunks store
  // their (discarded) outputs where the segment's first chunk later stores
  // its own, so every store is unconditional and one loop serves both.
  const int64_t M = a.M;
  C.load_next(cbot);
  C.park();
  __syncthreads();
  const int nwarm = a.warm / a.M;
  const int64_t first = cbot + a.warm;
  for (int k = 0; k < a.nchunks; ++k) {
    C.c = cbot + int64_t(k) * M;
    C.load_next(C.c + M);
    C.cst = k < nwarm ? first : C.c;
    C.template levels<G>();
    C.template shift_and_park<1>();
    __syncthreads();
  }
}

template <typename T, int FILT, int R, int NT, int G, int MAP>
__global__ __launch_bounds__(NT) void modwt_inv_k(ModwtArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t nrb = (int64_t(1) << a.j0) / R;
  const ModwtArgs b = batch_slice<T>(a);
  if constexpr (MAP == 2) {
    static_assert(R == 1, "linear buffers serve the first group");
    modwt_inv_lin_body<T, FILT, NT, G>(b, int64_t(blockIdx.x), smem_raw);
  } else {
    modwt_inv_body<T, FILT, R, NT, G, MAP>(b, int64_t(blockIdx.x) % nrb,
                                           int64_t(blockIdx.x) / nrb, smem_raw);
  }
}

// Launch of a (T, FILT) group kernel; defined by instantiation files.
struct GroupLaunch {
  int R, NT, G;
  int map;  // R = 1 element mapping: 0 adjacent pairs, 1 two outputs of one residue class (set by the planner), 2 the same with linear buffers (inverse)
  int sp;   // forward store policy (kStoreAll or kStoreAdjoint)
  int64_t blocks;
  int64_t count;  // signals (grid y)
  size_t lds;
};
template <typename T, int FILT>
int modwt_launch(bool inverse, const GroupLaunch& gl, const ModwtArgs& a, hipStream_t st);

}  // namespace wh
