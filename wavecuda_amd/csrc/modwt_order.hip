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
// A workgroup takes a tile of T consecutive time positions of one level:
// the S and D reads (or writes) of the tile are contiguous.  Through LDS
// the tile is regrouped by packet: with P = 2^{l+1} <= T packets the tile
// holds K = T / P consecutive pairs of every packet (runs of 16 K bytes);
// with more packets than positions it holds one pair of each of T packets.
// LDS rows are padded by one pair so the transposing accesses of
// consecutive lanes fall in different banks.
#include <algorithm>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {

namespace {

constexpr int kOrdT = 2048;  // time positions per tile
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
  const int64_t lenos = n >> l;
  const int64_t base = int64_t(blockIdx.x) << lg_tile;  // first time position of the tile
  const T* src = in + 2 * n * int64_t(l);
  T* dst = out + 2 * n * int64_t(l);
  const bool runs = P <= tile;
  const int lgK = runs ? lg_tile - bits : 0;  // K = 2^lgK pairs of each packet
  const int K = 1 << lgK;
  const int stride = K > 1 ? K + 1 : K;      // padded LDS row of one packet
  const int64_t k0 = base >> bits;           // first pair index in the tile
  const uint32_t r0 = uint32_t(base & (P - 1));

  // slot of the tile's time position j (j < tile)
  auto slot_of_time = [&](int j) -> int {
    if (!runs) return j;
    const int r = j & int(P - 1), kk = j >> bits;
    return r * stride + kk;
  };
  // e-th pair in packet order: its slot, its packet (time residue) and pair index
  auto packet_pair = [&](int e, int& slot, uint32_t& r, int64_t& k) {
    if (runs) {
      r = uint32_t(e >> lgK);
      const int kk = e & (K - 1);
      slot = int(r) * stride + kk;
      k = k0 + kk;
    } else {
      r = r0 + uint32_t(e);
      slot = e;
      k = k0;
    }
  };

  if (TO2PO) {
    for (int j = threadIdx.x; j < tile; j += kOrdNT) {
      const int64_t i = base + j;
      buf[slot_of_time(j)] = V{src[i], src[n + i]};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < tile; e += kOrdNT) {
      int slot;
      uint32_t r;
      int64_t k;
      packet_pair(e, slot, r, k);
      const int64_t p = int64_t(rev_bits(r, bits)) * lenos + 2 * k;
      *reinterpret_cast<V*>(dst + p) = buf[slot];
    }
  } else {
    for (int e = threadIdx.x; e < tile; e += kOrdNT) {
      int slot;
      uint32_t r;
      int64_t k;
      packet_pair(e, slot, r, k);
      const int64_t p = int64_t(rev_bits(r, bits)) * lenos + 2 * k;
      buf[slot] = *reinterpret_cast<const V*>(src + p);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < tile; j += kOrdNT) {
      const int64_t i = base + j;
      const V v = buf[slot_of_time(j)];
      dst[i] = v.x;
      dst[n + i] = v.y;
    }
  }
}

template <typename T>
int order_t(const void* in, void* out, int64_t n, int J, bool to_po, hipStream_t st) {
  int lg_tile = 0;
  while ((int64_t(1) << (lg_tile + 1)) <= std::min<int64_t>(n, kOrdT)) ++lg_tile;
  const int64_t tile = int64_t(1) << lg_tile;
  // padded rows: at most tile + tile / 2 pairs (K >= 2 means P <= tile / 2)
  const size_t lds = size_t(tile + tile / 2 + 1) * 2 * sizeof(T);
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
