// adjoint.hip -- adjoints for differentiating through the transforms and
// the soft threshold (gfx950).
//
// MODWT forward F: x -> (S_1, D_1, ..., S_J, D_J) with
//   S_j = H_j S_{j-1},  D_j = G_j S_{j-1},  (H_j u)[i] = sum_k C_k u[(i + k s) mod n]
// (reference src/daub4.cpp:561-571, s = 2^{j-1}).  Its adjoint takes the
// cotangents (Sb_j, Db_j) of every output back to x:
//   u_J = Sb_J,  u_{j-1} = Sb_{j-1} + H_j^T u_j + G_j^T Db_j,  xb = u_0
// with (H_j^T u)[p] = sum_k C_k u[(p - k s) mod n].  One kernel per level
// reads u_j, Db_j and Sb_{j-1} and writes u_{j-1}; each of its L taps is a
// shifted contiguous read, so the accesses are coalesced.  (The adjoint of
// the inverse is the fused forward with exact 2^-j scalings, modwt.hip; the
// DWT is orthogonal, so its forward and inverse are each other's adjoints.)
//
// Soft threshold y = sign(c) max(|c| - t, 0) on the details of levels
// [minlevel, maxlevel] (reference src/thresh.cpp:142-157): dy/dc = 1[|c| > t],
// dy/dt = -sign(c) 1[|c| > t] (0 at |c| = t, where the rule gives 0); the
// hard rule keeps c when |c| >= t: dy/dc = 1[|c| >= t], dy/dt = 0 almost
// everywhere.  The per-level sums of dy/dt * ybar are block partials reduced
// in a fixed order, so repeated runs give identical gradients.
#include <algorithm>

#include "wh_common.h"
#include "wh_internal.h"
#include "wh_stats.h"

namespace wh {

namespace {

template <typename T, int FILT>
__global__ __launch_bounds__(256) void modwt_fwd_adj_level_k(const T* __restrict__ u0,
                                                             int64_t bs_u,
                                                             const T* __restrict__ gxm0,
                                                             int64_t per, int j, T* __restrict__ out0,
                                                             int64_t n) {
  constexpr int L = Filter<FILT>::L;
  const int64_t sig = blockIdx.y;
  const T* __restrict__ u = u0 + sig * bs_u;
  const T* __restrict__ db = gxm0 + sig * per + 2 * n * int64_t(j - 1) + n;
  const T* __restrict__ sb = (j >= 2) ? gxm0 + sig * per + 2 * n * int64_t(j - 2) : nullptr;
  T* __restrict__ out = out0 + sig * n;
  const int64_t s = int64_t(1) << (j - 1);
  const int64_t mask = n - 1;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n;
       p += int64_t(gridDim.x) * blockDim.x) {
    T acc = sb ? sb[p] : T(0);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int64_t q = (p - k * s) & mask;
      acc = fma(u[q], T(Filter<FILT>::c(k)), acc);
      acc = fma(db[q], T(gcoef<FILT>(k)), acc);
    }
    out[p] = acc;
  }
}

template <typename T>
int fwd_adj_t(const void* gxmod, void* gx, int64_t n, int64_t count, int J, int filt,
              hipStream_t st) {
  const int64_t per = 2 * n * J;
  T* scratch = nullptr;
  if (J > 1) {
    void* p = nullptr;
    int rc = scratch_get(2 * size_t(n) * size_t(count) * sizeof(T), &p);
    if (rc) return rc;
    scratch = static_cast<T*>(p);
  }
  const T* gxm = static_cast<const T*>(gxmod);
  const T* u = gxm + 2 * n * int64_t(J - 1);  // u_J = Sb_J
  int64_t bs_u = per;
  const unsigned blocks = unsigned(std::min<int64_t>((n + 255) / 256, 2048));
  const dim3 grid(blocks, unsigned(count));
  for (int j = J; j >= 1; --j) {
    T* dst = (j == 1) ? static_cast<T*>(gx) : scratch + int64_t(j & 1) * n * count;
    switch (filt) {
      case kHAAR:
        hipLaunchKernelGGL((modwt_fwd_adj_level_k<T, kHAAR>), grid, dim3(256), 0, st, u, bs_u,
                           gxm, per, j, dst, n);
        break;
      case kD4:
        hipLaunchKernelGGL((modwt_fwd_adj_level_k<T, kD4>), grid, dim3(256), 0, st, u, bs_u, gxm,
                           per, j, dst, n);
        break;
      case kC6:
        hipLaunchKernelGGL((modwt_fwd_adj_level_k<T, kC6>), grid, dim3(256), 0, st, u, bs_u, gxm,
                           per, j, dst, n);
        break;
      default:
        hipLaunchKernelGGL((modwt_fwd_adj_level_k<T, kLA8>), grid, dim3(256), 0, st, u, bs_u,
                           gxm, per, j, dst, n);
        break;
    }
    WH_HIP_CHECK(hipGetLastError());
    u = dst;
    bs_u = n;
  }
  return WH_OK;
}

constexpr int kGradNT = 256;
constexpr int kGradMinPer = 4096;         // elements per workgroup, at least
constexpr int64_t kGradMaxPartials = 65536;  // partial sums per call, at most

template <typename T>
__device__ __forceinline__ void grad_one(T cv, T g, T t, int hardness, T* gcp, double& acc) {
  const T a = fabs(cv);
  if (hardness == kSOFT) {
    const bool pass = a > t;
    if (gcp) *gcp = pass ? g : T(0);
    if (pass) acc += double(cv < T(0) ? g : -g);
  } else if (gcp) {
    *gcp = (a < t) ? T(0) : g;
  }
}

// One workgroup per (chunk, level, signal).  MODWT details of a level are
// contiguous and move as element pairs; the DWT's are strided.  The
// workgroup's sum is reduced in a fixed order (lanes by shuffles, then the
// four waves), so results repeat bit for bit.
template <typename T>
__global__ __launch_bounds__(kGradNT) void thresh_grad_k(const T* __restrict__ c0,
                                                         const T* __restrict__ gy0,
                                                         T* __restrict__ gc0,
                                                         double* __restrict__ partial, int64_t n,
                                                         int J, int ttype, Thr1 th, int write_gc) {
  using V = typename Vec2<T>::type;
  const int l = th.minlevel + int(blockIdx.y);
  const int nl = int(gridDim.y);
  const int64_t sig = blockIdx.z;
  const int64_t per = (ttype == kDWT) ? n : 2 * n * J;
  const T* __restrict__ c = c0 + sig * per;
  const T* __restrict__ gy = gy0 + sig * per;
  T* __restrict__ gc = write_gc ? gc0 + sig * per : nullptr;
  const T t = th.levels ? T(th.levels[l]) : T(th.value);
  double acc = 0.0;
  if (ttype == kDWT) {
    const int64_t off = int64_t(1) << l, stride = int64_t(2) << l, cnt = n >> (l + 1);
    const int64_t chunk = (cnt + gridDim.x - 1) / gridDim.x;
    const int64_t e0 = int64_t(blockIdx.x) * chunk, e1 = std::min<int64_t>(cnt, e0 + chunk);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kGradNT) {
      const int64_t i = off + e * stride;
      grad_one(c[i], gy[i], t, th.hardness, gc ? gc + i : nullptr, acc);
    }
  } else {
    const int64_t off = 2 * n * l + n, cnt = n / 2;  // element pairs of D_{l+1}
    const int64_t chunk = (cnt + gridDim.x - 1) / gridDim.x;
    const int64_t e0 = int64_t(blockIdx.x) * chunk, e1 = std::min<int64_t>(cnt, e0 + chunk);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kGradNT) {
      const int64_t i = off + 2 * e;
      const V cv = *reinterpret_cast<const V*>(c + i);
      const V g = *reinterpret_cast<const V*>(gy + i);
      if (gc) {
        V o;
        grad_one(cv.x, g.x, t, th.hardness, &o.x, acc);
        grad_one(cv.y, g.y, t, th.hardness, &o.y, acc);
        *reinterpret_cast<V*>(gc + i) = o;
      } else {
        grad_one(cv.x, g.x, t, th.hardness, static_cast<T*>(nullptr), acc);
        grad_one(cv.y, g.y, t, th.hardness, static_cast<T*>(nullptr), acc);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  __shared__ double red[kGradNT / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[(sig * nl + (l - th.minlevel)) * int64_t(gridDim.x) + blockIdx.x] =
        (red[0] + red[1]) + (red[2] + red[3]);
}

// glam[l] = sum over signals and chunks of the level's partials (fixed
// order); zero outside [minlevel, maxlevel]
__global__ __launch_bounds__(kGradNT) void thresh_grad_reduce_k(const double* __restrict__ partial,
                                                                double* __restrict__ glam,
                                                                int64_t count, int nchunk,
                                                                int minlevel, int maxlevel) {
  const int l = blockIdx.x;
  __shared__ double red[kGradNT];
  double acc = 0.0;
  if (l >= minlevel && l <= maxlevel) {
    const int nl = maxlevel - minlevel + 1;
    const int64_t items = count * nchunk;
    for (int64_t k = threadIdx.x; k < items; k += kGradNT) {
      const int64_t sig = k / nchunk, ch = k % nchunk;
      acc += partial[(sig * nl + (l - minlevel)) * nchunk + ch];
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kGradNT / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) glam[l] = red[0];
}

}  // namespace

int modwt_forward_adjoint(const void* gxmod, void* gx, int64_t n, int64_t count, int J, int filt,
                          int dtype, hipStream_t st) {
  if (count < 1 || count > 65535) return set_error(WH_ERR_ARG, "adjoint: signal count out of range");
  return dtype == WH_F32 ? fwd_adj_t<float>(gxmod, gx, n, count, J, filt, st)
                         : fwd_adj_t<double>(gxmod, gx, n, count, J, filt, st);
}

int thresh_grad(const void* c, const void* gy, void* gc, double* glam, int64_t n, int64_t count,
                int J, int ttype, int dtype, const wh_thresh* thr, hipStream_t st) {
  if (count < 1 || count > 65535) return set_error(WH_ERR_ARG, "threshold gradient: signal count out of range");
  const size_t esz = dtype == WH_F32 ? 4 : 8;
  const int64_t per = (ttype == kDWT) ? n : 2 * n * J;
  // gc takes gy where the threshold does not act (scaling coefficients,
  // levels outside the range); without gc only glam is computed
  if (gc && gc != gy)
    WH_HIP_CHECK(hipMemcpyAsync(gc, gy, size_t(per) * size_t(count) * esz,
                                hipMemcpyDeviceToDevice, st));
  const int nl = thr->maxlevel - thr->minlevel + 1;
  // workgroups per (signal, level): at least kGradMinPer elements each, and
  // at most kGradMaxPartials partial sums in all
  const int64_t longest = ttype == kDWT ? n / 2 : n;
  const int64_t nchunk = std::max<int64_t>(
      1, std::min<int64_t>(longest / kGradMinPer, kGradMaxPartials / (count * nl)));
  void* p = nullptr;
  const size_t np = size_t(count) * size_t(nl) * size_t(nchunk);
  int rc = stats_get(np * sizeof(double), &p);
  if (rc) return rc;
  double* partial = static_cast<double*>(p);
  Thr1 th{thr->hardness, thr->minlevel, thr->maxlevel, thr->value, thr->levels};
  const dim3 grid{unsigned(nchunk), unsigned(nl), unsigned(count)};
  if (dtype == WH_F32)
    hipLaunchKernelGGL(thresh_grad_k<float>, grid, dim3(kGradNT), 0, st,
                       static_cast<const float*>(c), static_cast<const float*>(gy),
                       static_cast<float*>(gc), partial, n, J, ttype, th, gc ? 1 : 0);
  else
    hipLaunchKernelGGL(thresh_grad_k<double>, grid, dim3(kGradNT), 0, st,
                       static_cast<const double*>(c), static_cast<const double*>(gy),
                       static_cast<double*>(gc), partial, n, J, ttype, th, gc ? 1 : 0);
  WH_HIP_CHECK(hipGetLastError());
  if (glam) {
    hipLaunchKernelGGL(thresh_grad_reduce_k, dim3(unsigned(J)), dim3(kGradNT), 0, st, partial,
                       glam, count, int(nchunk), thr->minlevel, thr->maxlevel);
    WH_HIP_CHECK(hipGetLastError());
  }
  return WH_OK;
}

}  // namespace wh
