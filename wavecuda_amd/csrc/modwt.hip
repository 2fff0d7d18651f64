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
// level l has stride 2^{l-1} in m.  A workgroup owns R consecutive residues
// (R = min(B, 32), giving 8R-byte contiguous HBM runs) and one segment
// [mlo, mhi) of m, and marches through the segment in chunks of M
// (descending for the forward, whose taps look ahead; ascending for the
// inverse, whose taps look back).  Each level keeps an LDS ring holding the
// chunk plus the (L-1) 2^{l-1} window values left over from the previous
// chunk, so every level of the group reads its input from LDS and HBM sees
// each group input once and each output once.  The segment is entered
// after a warm-up of (L-1)(2^g - 1) m-positions (rounded up to M) whose
// outputs are computed but not stored -- the only redundant work.
// LDS layout is [slot][residue] so both the stride-2^{l-1} taps and the
// HBM runs are contiguous across lanes (no bank conflicts for R = 1 or 32).
// Loads of the next chunk are issued before a level's arithmetic and
// written to LDS after the level's barrier, hiding HBM latency behind
// compute with one workgroup per CU.
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
  int j0, g, logR, M;
  int warm;              // warm-up in m units, multiple of M
  int nchunks;           // (P + warm) / M
  // fused threshold (inverse only)
  int th_on, hardness, minlevel, maxlevel;
  double th_value;
  const double* th_levels;
};

// ---------------------------------------------------------------- forward
template <typename T, int FILT, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void modwt_fwd_group(ModwtArgs a) {
  constexpr int L = Filter<FILT>::L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);

  const int tid = threadIdx.x;
  const int logR = a.logR, R = 1 << logR, M = a.M, MR = M << logR;
  const int64_t n = a.n, Nm = n >> a.j0;
  const int64_t nrb = (int64_t(1) << a.j0) >> logR;
  const int64_t rb = int64_t(blockIdx.x) % nrb, seg = int64_t(blockIdx.x) / nrb;
  const int64_t r0 = rb << logR;
  const int64_t mlo = seg * a.P, mhi = mlo + a.P;
  const int g = a.g, j0 = a.j0;
  const T* __restrict__ in = static_cast<const T*>(a.in);
  T* __restrict__ xmod = static_cast<T*>(a.xmod);

  // first chunk start (descending march)
  const int64_t ctop = mhi + a.warm - M;

  // prologue: chunk 0 into ring 0 (head 0)
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    int q = tid + it * NT;
    if (q >= MR) break;
    int mm = q >> logR, r = q & (R - 1);
    int64_t m = pmod(ctop + mm, Nm);
    lds[mm * R + r] = in[r0 + r + (m << j0)];
  }
  __syncthreads();

  for (int k = 0; k < a.nchunks; ++k) {
    const int64_t c = ctop - int64_t(k) * M;
    const bool have_next = (k + 1 < a.nchunks);
    T pre[ITEMS];
    for (int l = 1; l <= g; ++l) {
      const int s = 1 << (l - 1);
      const int Qin = M + (L - 1) * s;
      const int offin = ((l - 1) * M + (L - 1) * (s - 1)) * R;
      const int headin = (Qin - int((int64_t(k) * M) % Qin)) % Qin;
      const int Qout = M + (L - 1) * 2 * s;
      const int offout = (l * M + (L - 1) * (2 * s - 1)) * R;
      const int headout = (Qout - int((int64_t(k) * M) % Qout)) % Qout;
      const bool store = (l < g);
      T* __restrict__ oS = xmod + 2 * n * int64_t(j0 + l - 1);
      T* __restrict__ oD = oS + n;

      if (l == 1 && have_next) {
        const int64_t cn = c - M;
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
          int q = tid + it * NT;
          if (q >= MR) break;
          int mm = q >> logR, r = q & (R - 1);
          int64_t m = pmod(cn + mm, Nm);
          pre[it] = in[r0 + r + (m << j0)];
        }
      }

#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        int q = tid + it * NT;
        if (q >= MR) break;
        int mm = q >> logR, r = q & (R - 1);
        int slot = headin + mm;
        if (slot >= Qin) slot -= Qin;
        T sv, dv;
        if constexpr (FILT == kHAAR) {
          T v0 = lds[offin + slot * R + r];
          int s1 = slot + s;
          if (s1 >= Qin) s1 -= Qin;
          T v1 = lds[offin + s1 * R + r];
          sv = (v0 + v1) * T(kInvR2);
          dv = (v0 - v1) * T(kInvR2);
        } else {
          T v = lds[offin + slot * R + r];
          sv = v * T(Filter<FILT>::c(0));
          dv = v * T(gcoef<FILT>(0));
#pragma unroll
          for (int kk = 1; kk < L; ++kk) {
            slot += s;
            if (slot >= Qin) slot -= Qin;
            v = lds[offin + slot * R + r];
            sv = fma(v, T(Filter<FILT>::c(kk)), sv);
            dv = fma(v, T(gcoef<FILT>(kk)), dv);
          }
        }
        if (store) {
          int so = headout + mm;
          if (so >= Qout) so -= Qout;
          lds[offout + so * R + r] = sv;
        }
        const int64_t m = c + mm;
        if (m >= mlo && m < mhi) {
          const int64_t i = r0 + r + (m << j0);
          __builtin_nontemporal_store(sv, oS + i);
          __builtin_nontemporal_store(dv, oD + i);
        }
      }
      __syncthreads();
      if (l == 1 && have_next) {
        // ring 0 slots of chunk k+1: head0(k+1) + mm
        const int Q0 = M + (L - 1);
        const int h0 = (Q0 - int((int64_t(k + 1) * M) % Q0)) % Q0;
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
          int q = tid + it * NT;
          if (q >= MR) break;
          int mm = q >> logR, r = q & (R - 1);
          int so = h0 + mm;
          if (so >= Q0) so -= Q0;
          lds[so * R + r] = pre[it];
        }
        if (g == 1) __syncthreads();
      }
    }
  }
}

// ---------------------------------------------------------------- inverse
template <typename T, int FILT, int NT, int ITEMS>
__global__ __launch_bounds__(NT) void modwt_inv_group(ModwtArgs a) {
  constexpr int L = Filter<FILT>::L;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);

  const int tid = threadIdx.x;
  const int logR = a.logR, R = 1 << logR, M = a.M, MR = M << logR;
  const int64_t n = a.n, Nm = n >> a.j0;
  const int64_t nrb = (int64_t(1) << a.j0) >> logR;
  const int64_t rb = int64_t(blockIdx.x) % nrb, seg = int64_t(blockIdx.x) / nrb;
  const int64_t r0 = rb << logR;
  const int64_t mlo = seg * a.P, mhi = mlo + a.P;
  const int g = a.g, j0 = a.j0;
  const T* __restrict__ xmod = static_cast<const T*>(a.xmod);
  const T* __restrict__ stop = static_cast<const T*>(a.stop);
  T* __restrict__ out = static_cast<T*>(a.out);
  // ring l (1..g) size Q_l = M + W_l, W_l = (L-1) 2^{l-1}; S rings first,
  // then D rings.
  const int baseD = (g * M + (L - 1) * ((1 << g) - 1)) * R;
  const int64_t cbot = mlo - a.warm;  // first chunk start (ascending march)

  auto thr_of = [&](int l) -> T {  // threshold of group level l (1-based)
    const int lev = j0 + l - 1;
    if (!a.th_on || lev < a.minlevel || lev > a.maxlevel) return T(-1);
    return a.th_levels ? T(a.th_levels[lev]) : T(a.th_value);
  };

  // prologue: every ring's chunk-0 region (offsets W_l + mm, head 0)
  for (int l = 1; l <= g; ++l) {
    const int W = (L - 1) << (l - 1);
    const int offS = ((l - 1) * M + (L - 1) * ((1 << (l - 1)) - 1)) * R;
    const T* __restrict__ dsrc = xmod + 2 * n * int64_t(j0 + l - 1) + n;
    const T t = thr_of(l);
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      int q = tid + it * NT;
      if (q >= MR) break;
      int mm = q >> logR, r = q & (R - 1);
      int64_t m = pmod(cbot + mm, Nm);
      int64_t i = r0 + r + (m << j0);
      T d = dsrc[i];
      if (t >= T(0)) d = thresh_coef(d, t, a.hardness);
      lds[baseD + offS + (W + mm) * R + r] = d;
      if (l == g) lds[offS + (W + mm) * R + r] = stop[i];
    }
  }
  __syncthreads();

  for (int k = 0; k < a.nchunks; ++k) {
    const int64_t c = cbot + int64_t(k) * M;
    const bool have_next = (k + 1 < a.nchunks);
    for (int l = g; l >= 1; --l) {
      const int s = 1 << (l - 1);
      const int W = (L - 1) * s;
      const int Q = M + W;
      const int offS = ((l - 1) * M + (L - 1) * (s - 1)) * R;
      const int head = int((int64_t(k) * M) % Q);
      T dpre[ITEMS], spre[ITEMS];
      const T t = thr_of(l);
      const T* __restrict__ dsrc = xmod + 2 * n * int64_t(j0 + l - 1) + n;
      if (have_next) {
        const int64_t cn = c + M;
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
          int q = tid + it * NT;
          if (q >= MR) break;
          int mm = q >> logR, r = q & (R - 1);
          int64_t m = pmod(cn + mm, Nm);
          int64_t i = r0 + r + (m << j0);
          dpre[it] = dsrc[i];
          if (l == g) spre[it] = stop[i];
        }
      }
      // output ring (level l-1): Q_{l-1} = M + W/2, head_{l-1}
      const int Wo = (l > 1) ? (L - 1) * (s >> 1) : 0;
      const int Qo = M + Wo;
      const int offSo = (l > 1) ? ((l - 2) * M + (L - 1) * ((s >> 1) - 1)) * R : 0;
      const int heado = (l > 1) ? int((int64_t(k) * M) % Qo) : 0;
#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        int q = tid + it * NT;
        if (q >= MR) break;
        int mm = q >> logR, r = q & (R - 1);
        // position m = c + mm sits at offset W + mm; tap k at offset W + mm - k s
        T acc;
        if constexpr (FILT == kHAAR) {
          int s0 = head + W + mm;
          if (s0 >= Q) s0 -= Q;
          int s1 = s0 - s;
          if (s1 < 0) s1 += Q;
          T S0 = lds[offS + s0 * R + r], D0 = lds[baseD + offS + s0 * R + r];
          T S1 = lds[offS + s1 * R + r], D1 = lds[baseD + offS + s1 * R + r];
          acc = (S0 + D0 + S1 - D1) * T(kHInvR2);
        } else {
          // term order k = L-2, L-4, ..., 0, L-1, ..., 1 (src/daub4.cpp:578-583)
          acc = T(0);
          bool first = true;
#pragma unroll
          for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int kk = L - 2 + pass; kk >= 0; kk -= 2) {
              int so = head + W + mm - kk * s;
              if (so >= Q) so -= Q;
              if (so < 0) so += Q;
              T Sv = lds[offS + so * R + r], Dv = lds[baseD + offS + so * R + r];
              if (first) {
                acc = Sv * T(Filter<FILT>::c(kk));
                first = false;
              } else {
                acc = fma(Sv, T(Filter<FILT>::c(kk)), acc);
              }
              acc = fma(Dv, T(gcoef<FILT>(kk)), acc);
            }
          }
          acc = T(0.5) * acc;
        }
        if (l > 1) {
          int so = heado + Wo + mm;
          if (so >= Qo) so -= Qo;
          lds[offSo + so * R + r] = acc;
        } else {
          const int64_t m = c + mm;
          if (m >= mlo && m < mhi) __builtin_nontemporal_store(acc, out + r0 + r + (m << j0));
        }
      }
      __syncthreads();
      if (have_next) {
        // chunk k+1 region of ring l: head(k+1) + W + mm
        const int hn = int((int64_t(k + 1) * M) % Q);
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
          int q = tid + it * NT;
          if (q >= MR) break;
          int mm = q >> logR, r = q & (R - 1);
          int so = hn + W + mm;
          if (so >= Q) so -= Q;
          T d = dpre[it];
          if (t >= T(0)) d = thresh_coef(d, t, a.hardness);
          lds[baseD + offS + so * R + r] = d;
          if (l == g) lds[offS + so * R + r] = spre[it];
        }
        if (g == 1) __syncthreads();
      }
    }
  }
}

// ------------------------------------------------------------------ host

namespace {

constexpr int kNT = 256;
constexpr int kRmax = 32;

// positions per thread per chunk: the inverse keeps two rings per level
// (S and D), so it runs half-size chunks to fit 10 fused LA8 levels in LDS
template <typename T, bool INV>
constexpr int items_for() {
  return (sizeof(T) == 8 ? 2 : 4) / (INV ? 2 : 1);
}

struct Group {
  int j0, g, logR, M;
  int64_t P;
  int nseg, warm, nchunks;
  size_t lds;
};

size_t lds_bytes(bool inverse, int L, int g, int M, int R, size_t esz) {
  size_t per = size_t(g) * M + size_t(L - 1) * ((size_t(1) << g) - 1);
  return (inverse ? 2 : 1) * per * R * esz;
}

int64_t round_up64(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Partition levels 1..J into groups; per group choose residues per tile,
// chunk length, fused depth and segment count.  Depth is limited by the
// LDS budget and by keeping the warm-up at most half a segment when the
// grid fills every CU.
std::vector<Group> plan_groups(bool inverse, int64_t n, int J, int L, size_t esz, int items) {
  const DeviceInfo& dev = device_info();
  const size_t budget = std::min<size_t>(dev.max_lds_per_block, 160 * 1024);
  const int64_t CH = int64_t(kNT) * items;  // positions per chunk
  const int64_t cus = std::max(1, dev.num_cus);
  std::vector<Group> gs;
  int j0 = 0;
  while (j0 < J) {
    Group G{};
    G.j0 = j0;
    const int64_t B = int64_t(1) << j0;
    const int R = int(std::min<int64_t>(B, kRmax));
    G.logR = 0;
    while ((1 << G.logR) < R) ++G.logR;
    const int64_t Nm = n >> j0;
    G.M = int(std::min<int64_t>(CH / R, Nm));
    const int64_t nrb = B / R;
    const int64_t target = std::max<int64_t>(1, cus / nrb);
    int64_t nseg_t = 1;
    while (nseg_t * 2 <= target && Nm / (nseg_t * 2) >= G.M) nseg_t *= 2;
    const int64_t Pt = Nm / nseg_t;
    int gbest = 1;
    for (int g = 2; g <= J - j0; ++g) {
      if (lds_bytes(inverse, L, g, G.M, R, esz) > budget) break;
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
    G.lds = lds_bytes(inverse, L, G.g, G.M, R, esz);
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

template <typename T, int FILT>
int launch_fwd(const Group& G, const ModwtArgs& a, hipStream_t st) {
  constexpr int IT = items_for<T, false>();
  auto kern = modwt_fwd_group<T, FILT, kNT, IT>;
  static std::once_flag once;
  std::call_once(once, [&] { allow_big_lds(kern); });
  int64_t nrb = (int64_t(1) << G.j0) >> G.logR;
  int64_t blocks = nrb * G.nseg;
  if (blocks > 0x7fffffff) return set_error(WH_ERR_ARG, "modwt: grid too large");
  hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kNT), G.lds, st, a);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

template <typename T, int FILT>
int launch_inv(const Group& G, const ModwtArgs& a, hipStream_t st) {
  constexpr int IT = items_for<T, true>();
  auto kern = modwt_inv_group<T, FILT, kNT, IT>;
  static std::once_flag once;
  std::call_once(once, [&] { allow_big_lds(kern); });
  int64_t nrb = (int64_t(1) << G.j0) >> G.logR;
  int64_t blocks = nrb * G.nseg;
  if (blocks > 0x7fffffff) return set_error(WH_ERR_ARG, "modwt: grid too large");
  hipLaunchKernelGGL(kern, dim3(unsigned(blocks)), dim3(kNT), G.lds, st, a);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

ModwtArgs args_of(const Group& G, int64_t n) {
  ModwtArgs a{};
  a.n = n;
  a.P = G.P;
  a.j0 = G.j0;
  a.g = G.g;
  a.logR = G.logR;
  a.M = G.M;
  a.warm = G.warm;
  a.nchunks = G.nchunks;
  return a;
}

template <typename T>
int modwt_fwd_t(const void* x, void* xmod, int64_t n, int J, int filt, hipStream_t st) {
  const int L = filter_length(filt);
  auto groups = plan_groups(false, n, J, L, sizeof(T), items_for<T, false>());
  for (const Group& G : groups) {
    ModwtArgs a = args_of(G, n);
    a.in = (G.j0 == 0) ? x : static_cast<const T*>(xmod) + 2 * n * int64_t(G.j0 - 1);
    a.xmod = xmod;
    int rc;
    switch (filt) {
      case kHAAR: rc = launch_fwd<T, kHAAR>(G, a, st); break;
      case kD4: rc = launch_fwd<T, kD4>(G, a, st); break;
      case kC6: rc = launch_fwd<T, kC6>(G, a, st); break;
      default: rc = launch_fwd<T, kLA8>(G, a, st); break;
    }
    if (rc) return rc;
  }
  return WH_OK;
}

template <typename T>
int modwt_inv_t(const void* xmod, void* x, int64_t n, int J, int filt, const wh_thresh* thr,
                hipStream_t st) {
  const int L = filter_length(filt);
  auto groups = plan_groups(true, n, J, L, sizeof(T), items_for<T, true>());
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
    int rc;
    switch (filt) {
      case kHAAR: rc = launch_inv<T, kHAAR>(G, a, st); break;
      case kD4: rc = launch_inv<T, kD4>(G, a, st); break;
      case kC6: rc = launch_inv<T, kC6>(G, a, st); break;
      default: rc = launch_inv<T, kLA8>(G, a, st); break;
    }
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
  const bool f32 = dtype == WH_F32;
  size_t esz = f32 ? 4 : 8;
  int items = f32 ? (inverse ? items_for<float, true>() : items_for<float, false>())
                  : (inverse ? items_for<double, true>() : items_for<double, false>());
  auto gs = plan_groups(inverse, n, J, L, esz, items);
  std::string s;
  for (auto& G : gs) {
    char buf[256];
    snprintf(buf, sizeof buf,
             "{j0=%d g=%d R=%d M=%d nseg=%d P=%lld warm=%d chunks=%d lds=%zu} ", G.j0, G.g,
             1 << G.logR, G.M, G.nseg, (long long)G.P, G.warm, G.nchunks, G.lds);
    s += buf;
  }
  return s;
}

}  // namespace wh
