// Device-side helpers shared by the libhmotif kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hm {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kStaticBins = 30;
constexpr int kOpenBin0 = 20;               // bins 20..25: open triples (MoCHy printed motifs 17-22)
// Packed projection-row entry: (neighbour << kCntBits) | overlap.
constexpr int kCntBits = 7;
constexpr uint32_t kCntMask = (1u << kCntBits) - 1u;
constexpr int64_t kMaxPackedId = (int64_t)1 << (32 - kCntBits);

// 64-bit masks use unsigned long long (the type of the 64-bit LDS atomics).
using mask64_t = unsigned long long;
__device__ inline int popc(uint32_t m) { return __popc(m); }
__device__ inline int popc(mask64_t m) { return __popcll(m); }

__device__ inline int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ inline int wave_id() { return (int)(threadIdx.x / kWave); }

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline unsigned long long wave_max(unsigned long long v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

// Adds `delta` to h[bin] for every lane with `valid`, aggregating equal bins across the wave:
// each round takes the first remaining lane's bin, counts the lanes that share it with one
// ballot, and lets that lane update the counter. Rounds = distinct bins in the wave. `h` must be
// private to the calling wave (a plain read-modify-write by one lane at a time).
template <typename CT>
__device__ inline void wave_hist_add(CT* h, int bin, bool valid, CT delta) {
  unsigned long long rem = __ballot(valid);
  while (rem) {
    const int leader = __ffsll((long long)rem) - 1;
    const int b0 = __shfl(bin, leader);
    const unsigned long long m = __ballot(valid && bin == b0);
    if (lane_id() == leader) h[b0] += delta * (CT)__popcll(m);
    rem &= ~m;
  }
}

// Venn-region emptiness id of (x1, x2, x3) (MoCHy-master/motif_id.cpp:24-34):
// bit0..2 = x1/x2/x3-only regions, bit3 = x1x2-only, bit4 = x2x3-only, bit5 = x3x1-only, bit6 = all.
__device__ inline int venn_id(int s1, int s2, int s3, int c12, int c23, int c31, int g) {
  return (int)(s1 - c12 - c31 + g > 0) | ((int)(s2 - c12 - c23 + g > 0) << 1) |
         ((int)(s3 - c31 - c23 + g > 0) << 2) | ((int)(c12 > g) << 3) | ((int)(c23 > g) << 4) |
         ((int)(c31 > g) << 5) | ((int)(g > 0) << 6);
}

// Bin of an open triple centred at x (x overlaps y and z, y and z disjoint):
// MoCHy bins 20..25 = 20 + [x has a private node] + 2 * ([y ⊄ x] + [z ⊄ x]).
__device__ inline int open_bin(int sx, int sy, int sz, int cxy, int cxz) {
  return kOpenBin0 + (int)(sx - cxy - cxz > 0) + 2 * ((int)(sy > cxy) + (int)(sz > cxz));
}

// Mask table: M[key(x)] |= bit(i) for every node a_i of hyperedge a and every x in V2H(a_i), x != a.
template <typename MT>
__device__ inline void build_masks(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                   const int32_t* __restrict__ enodes,
                                   const uint32_t* __restrict__ voff,
                                   const int32_t* __restrict__ vedges,
                                   const int32_t* __restrict__ rank) {
  for (int x = threadIdx.x; x < E; x += blockDim.x) M[x] = (MT)0;
  __syncthreads();
  const uint32_t s = eoff[a], t = eoff[a + 1];
  for (uint32_t i = s; i < t; ++i) {
    const int32_t v = enodes[i];
    const MT bit = (MT)1 << (i - s);
    const uint32_t ks = voff[v], kt = voff[v + 1];
    for (uint32_t k = ks + threadIdx.x; k < kt; k += blockDim.x) {
      const int32_t x = vedges[k];
      if (x != a) atomicOr(&M[rank ? rank[x] : x], bit);
    }
  }
  __syncthreads();
}

}  // namespace hm
