// levels.hip -- the coefficient layouts read by level (gfx950).
//
//  * gather of one level's detail or scaling coefficients into a compact
//    array: the positions the reference's getCoeffLevel lists
//    (R/rplotting.R:10-46: DWT details of level l at 2^{l-1} + m 2^l,
//    scaling coefficients at m 2^J) and, for a time-ordered MODWT, the
//    level's S_l or D_l block (xmod[2n(l-1) + i], xmod[2n(l-1) + n + i],
//    src/daub4.cpp:525-528);
//  * the packed per-level vectors the reference's WSTtoWavethresh hands to
//    the wavethresh package (R/rfunctions.R:864-893);
//  * the label columns the reference's WSTtoDT attaches to every
//    coefficient (R/rfunctions.R:925-1004) and label.detail.scaling
//    (:902-912): coefficient type, level and translate, one entry per
//    coefficient in storage order.
//
// Both are index arithmetic over at most 2 n J entries: memory-bound
// grid-stride loops, one element per thread and iteration.
#include <algorithm>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {

namespace {

constexpr int kLevNT = 256;

template <typename T>
__global__ void level_gather_k(const T* __restrict__ c, T* __restrict__ out, int64_t start,
                               int64_t gap, int64_t count) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < count; m += stride)
    out[m] = c[start + m * gap];
}

// DWT position i of n = 2^k coefficients with J levels: i % 2^J == 0 holds
// s_J[i / 2^J]; otherwise the lowest set bit b gives d_{b+1}[i >> (b+1)].
// translate of a level-l coefficient m: 2^{l-1} (m + 1/2), the value
// WSTtoDT assigns (R/rfunctions.R:944-970), which gives the scaling
// coefficients the translates of level J's details.
__global__ void labels_dwt_k(int64_t n, int J, int* __restrict__ level,
                             double* __restrict__ translate, unsigned char* __restrict__ type) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const uint64_t smask = (uint64_t(1) << J) - 1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const bool s = (uint64_t(i) & smask) == 0;
    const int l = s ? J : __builtin_ctzll(uint64_t(i)) + 1;
    const int64_t m = int64_t(uint64_t(i) >> l);
    if (level) level[i] = l;
    if (translate) translate[i] = double(int64_t(1) << (l - 1)) * (double(m) + 0.5);
    if (type) type[i] = s ? 1 : 0;
  }
}

// time-ordered MODWT: entry i of 2 n J is S (first n of each 2n block) or D
// of level i / 2n + 1, translate (i mod n) + 1 (R/rfunctions.R:979-985)
__global__ void labels_modwt_k(int64_t n, int J, int* __restrict__ level,
                               double* __restrict__ translate, unsigned char* __restrict__ type) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t total = 2 * n * int64_t(J);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int64_t q = i / n;
    if (level) level[i] = int(q / 2) + 1;
    if (translate) translate[i] = double(i - q * n + 1);
    if (type) type[i] = (q & 1) ? 0 : 1;
  }
}

unsigned grid_for(int64_t count) {
  const int64_t cus = std::max(1, device_info().num_cus);
  const int64_t want = (count + kLevNT - 1) / kLevNT;
  return unsigned(std::max<int64_t>(1, std::min<int64_t>(want, cus * 16)));
}

}  // namespace

int level_extract(const void* c, void* out, int64_t n, int J, int ttype, int level, int coeff_s,
                  int dtype, hipStream_t st, int64_t* count_out) {
  int64_t start, gap, count;
  if (ttype == kDWT) {
    start = coeff_s ? 0 : (int64_t(1) << (level - 1));
    gap = coeff_s ? (int64_t(1) << J) : (int64_t(1) << level);
    count = n >> (coeff_s ? J : level);
  } else {
    start = 2 * n * int64_t(level - 1) + (coeff_s ? 0 : n);
    gap = 1;
    count = n;
  }
  if (count_out) *count_out = count;
  const unsigned grid = grid_for(count);
  if (dtype == WH_F32)
    hipLaunchKernelGGL(level_gather_k<float>, dim3(grid), dim3(kLevNT), 0, st,
                       static_cast<const float*>(c), static_cast<float*>(out), start, gap, count);
  else
    hipLaunchKernelGGL(level_gather_k<double>, dim3(grid), dim3(kLevNT), 0, st,
                       static_cast<const double*>(c), static_cast<double*>(out), start, gap,
                       count);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

// WSTtoWavethresh (R/rfunctions.R:864-893): the vectors its loop hands to
// wavethresh's putD for l = 1..J (level log2(n) - l) and then to putC (the
// level-J scaling coefficients), packed back to back in that order
int wavethresh_pack(const void* c, void* out, int64_t n, int J, int ttype, int dtype,
                    hipStream_t st) {
  const size_t esz = dtype == WH_F32 ? 4 : 8;
  int64_t off = 0;
  for (int e = 1; e <= J + 1; ++e) {
    const int level = e <= J ? e : J;
    const int coeff_s = e <= J ? 0 : 1;
    int64_t cnt = 0;
    int rc = level_extract(c, static_cast<char*>(out) + size_t(off) * esz, n, J, ttype, level,
                           coeff_s, dtype, st, &cnt);
    if (rc) return rc;
    off += cnt;
  }
  return WH_OK;
}

int level_labels(int64_t n, int J, int ttype, int* level, double* translate, unsigned char* type,
                 hipStream_t st) {
  if (ttype == kDWT)
    hipLaunchKernelGGL(labels_dwt_k, dim3(grid_for(n)), dim3(kLevNT), 0, st, n, J, level,
                       translate, type);
  else
    hipLaunchKernelGGL(labels_modwt_k, dim3(grid_for(2 * n * int64_t(J))), dim3(kLevNT), 0, st,
                       n, J, level, translate, type);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

}  // namespace wh
