// modwt_order.hip -- time-ordered <-> packet-ordered MODWT coefficients (gfx950).
//
// Level l (0-based) of a packet-ordered MODWT holds 2^{l+1} packets of
// n / 2^l values, (s, d) pairs; packet `shift` carries the time positions
// rev(shift) + 2^{l+1} k, rev reversing the l + 1 low bits (reference
// src/wavutils.cpp:565-600, the reversal at :579-583).  The time-ordered
// layout keeps S_{l+1} and D_{l+1} as two runs of n values.  Converting is a
// permutation of each level's 2n values, so every value is read once and
// written once (algorithmic bytes: 16 per value in FP64, 8 in FP32).
//
// A workgroup takes a tile of T = A K time positions of one level: A
// consecutive time residues (packets rev(r), r in [ra A, ra A + A)) times K
// consecutive pair indices k of each.  The time-ordered side then moves K
// runs of A positions of S and of D, the packet-ordered side A runs of K
// (s, d) pairs; A = P = 2^{l+1} while P <= 32 (the tile is one contiguous
// time range), otherwise A = 32, K = T / 32, so both sides move runs of at
// least 256 bytes (FP64) at every level.  Through LDS the tile is held as
// A rows of K pairs, each row padded by one pair so the transposing
// accesses of consecutive lanes fall in different banks.  The time-ordered
// side moves two positions per 16-byte access.
#include <algorithm>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {

namespace {

constexpr int kOrdT = 2048;  // time positions per tile
constexpr int kOrdA = 32;    // residues per tile once the level has more packets
constexpr int kOrdNT = 256;

__device__ __forceinline__ uint32_t rev_bits(uint32_t v, int bits) {
  return __brev(v) >> (32 - bits);
}

// TO2PO: time ordered in, packet ordered out; otherwise the reverse
template <typename T, bool TO2PO>
__global__ __launch_bounds__(kOrdNT) void modwt_order_k(const T* __restrict__ in,
                                                        T* __restrict__ out, int64_t n,
                                                        int lg_tile) {
  using V = typename Vec2<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  V* buf = reinterpret_cast<V*>(smem);
  const int tile = 1 << lg_tile;
  const int l = blockIdx.y;
  const int bits = l + 1;
  const int64_t P = int64_t(1) << bits;
  const int64_t lenos = n >> l;          // values per packet
  const int64_t npairs = n >> bits;      // pairs per packet
  // K pairs of each of A residues: K = min(T / min(P, 32), n / P), A = T / K
  const int64_t A0 = P < kOrdA ? P : kOrdA;
  const int64_t K64 = std::min<int64_t>(tile / A0, npairs);
  const int lgK = 31 - __clz(int(K64));
  const int K = 1 << lgK;
  const int lgA = lg_tile - lgK;
  const int A = 1 << lgA;
  const int64_t nra = P >> lgA;          // residue blocks of the level
  const int64_t ra = int64_t(blockIdx.x) % nra, kb = int64_t(blockIdx.x) / nra;
  const int64_t r0 = ra << lgA;          // first residue of the tile
  const int64_t k0 = kb << lgK;          // first pair index of the tile
  const int stride = K > 1 ? K + 1 : K;  // padded LDS row of one residue
  const T* src = in + 2 * n * int64_t(l);
  T* dst = out + 2 * n * int64_t(l);
  const int half = A >> 1;               // position pairs per run of the time side

  if (TO2PO) {
    for (int q = threadIdx.x; q < tile / 2; q += kOrdNT) {
      const int a = 2 * (q & (half - 1)), kk = q >> (lgA - 1);
      const int64_t i = r0 + a + P * (k0 + kk);
      const V sv = *reinterpret_cast<const V*>(src + i);
      const V dv = *reinterpret_cast<const V*>(src + n + i);
      buf[a * stride + kk] = V{sv.x, dv.x};
      buf[(a + 1) * stride + kk] = V{sv.y, dv.y};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < tile; e += kOrdNT) {
      const int a = e >> lgK, kk = e & (K - 1);
      const int64_t p = int64_t(rev_bits(uint32_t(r0 + a), bits)) * lenos + 2 * (k0 + kk);
      *reinterpret_cast<V*>(dst + p) = buf[a * stride + kk];
    }
  } else {
    for (int e = threadIdx.x; e < tile; e += kOrdNT) {
      const int a = e >> lgK, kk = e & (K - 1);
      const int64_t p = int64_t(rev_bits(uint32_t(r0 + a), bits)) * lenos + 2 * (k0 + kk);
      buf[a * stride + kk] = *reinterpret_cast<const V*>(src + p);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < tile / 2; q += kOrdNT) {
      const int a = 2 * (q & (half - 1)), kk = q >> (lgA - 1);
      const int64_t i = r0 + a + P * (k0 + kk);
      const V v0 = buf[a * stride + kk], v1 = buf[(a + 1) * stride + kk];
      *reinterpret_cast<V*>(dst + i) = V{v0.x, v1.x};
      *reinterpret_cast<V*>(dst + n + i) = V{v0.y, v1.y};
    }
  }
}

template <typename T>
int order_t(const void* in, void* out, int64_t n, int J, bool to_po, hipStream_t st) {
  int lg_tile = 0;
  while ((int64_t(1) << (lg_tile + 1)) <= std::min<int64_t>(n, kOrdT)) ++lg_tile;
  const int64_t tile = int64_t(1) << lg_tile;
  if (tile < 2) return set_error(WH_ERR_ARG, "modwt convert: length must be at least 2");
  // padded rows: A (K + 1) = tile + A pairs, A <= tile / 2
  const size_t lds = size_t(tile + tile / 2) * 2 * sizeof(T);
  const dim3 grid(unsigned(n / tile), unsigned(J));
  if (to_po)
    hipLaunchKernelGGL((modwt_order_k<T, true>), grid, dim3(kOrdNT), lds, st,
                       static_cast<const T*>(in), static_cast<T*>(out), n, lg_tile);
  else
    hipLaunchKernelGGL((modwt_order_k<T, false>), grid, dim3(kOrdNT), lds, st,
                       static_cast<const T*>(in), static_cast<T*>(out), n, lg_tile);
  WH_HIP_CHECK(hipGetLastError());
  return WH_OK;
}

}  // namespace

int modwt_order(const void* in, void* out, int64_t n, int J, bool to_po, int dtype,
                hipStream_t st) {
  return dtype == WH_F32 ? order_t<float>(in, out, n, J, to_po, st)
                         : order_t<double>(in, out, n, J, to_po, st);
}

}  // namespace wh
