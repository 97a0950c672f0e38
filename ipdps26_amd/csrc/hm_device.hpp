// Device-side helpers shared by the libhmotif kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hmotif.h"

namespace hm {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kStaticBins = 30;
constexpr int kOpenBin0 = 20;               // bins 20..25: open triples (MoCHy printed motifs 17-22)
// Accumulator slots of one graph's count (uint64).
enum AccSlot : int {
  kAccHist = 0,        // 30 bins (two's-complement wrap: intermediate -1 contributions are exact)
  kAccP = 32,          // sum_a d_a
  kAccMaxDeg = 33,     // max_a d_a
  kAccOrient = 34,     // sum_a sum_{b in N(a), b > a} d_b
  kAccWedges = 35,     // sum_a C(d_a, 2)
  kAccUpper = 36,      // sum_a d+_a (upper-row entries)
  kAccMaxNodeDeg = 37, // max_v |V2H(v)|
  kAccLightTri = 38,   // light triangles (hub-aware count)
  kAccCount = 40
};

// Packed projection-row entry: (neighbour << kCntBits) | overlap.
constexpr int kCntBits = 7;
constexpr uint32_t kCntMask = (1u << kCntBits) - 1u;
constexpr int64_t kMaxPackedId = (int64_t)1 << (32 - kCntBits);
// Ids (hyperedges, temporal hyperedges) a graph may hold: the kernels launch one workgroup of up to
// 1,024 work-items per id, and a launch's work-item count (grid x block) is a 32-bit quantity.
constexpr int64_t kMaxIds = (int64_t)1 << 22;
static_assert(kMaxIds <= kMaxPackedId, "ids must fit the packed row entries");

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

// Buffer descriptor over n packed row entries at p (raw, no stride; word 3 as for gfx9 raw
// buffers): a read at entry index >= n returns 0. Indices passed to row_load must be
// non-negative and small enough that 4 * index does not wrap.
// p and n must be the same on every active lane (one row per wave); they are read from the
// first lane so that the descriptor sits in scalar registers.
__device__ inline __amdgpu_buffer_rsrc_t row_rsrc(const uint32_t* p, uint32_t n) {
  const unsigned long long a = (unsigned long long)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  const uint32_t nn = __builtin_amdgcn_readfirstlane(n);
  uint32_t* q = reinterpret_cast<uint32_t*>((unsigned long long)hi << 32 | lo);
  return __builtin_amdgcn_make_buffer_rsrc(q, (short)0, (int)(nn * 4u), 0x00020000);
}

__device__ inline uint32_t row_load(__amdgpu_buffer_rsrc_t rs, uint32_t j) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(j * 4u), 0, 0);
}

// Venn-region emptiness id of (x1, x2, x3) (MoCHy-master/motif_id.cpp:24-34):
// bit0..2 = x1/x2/x3-only regions, bit3 = x1x2-only, bit4 = x2x3-only, bit5 = x3x1-only, bit6 = all.
// [r > 0] of a region count r >= 0 (one unsigned min)
__device__ inline unsigned nz(int r) { return min((unsigned)r, 1u); }

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
// With `hub` given, nodes v with hub[v] >= 0 are skipped (the table then holds a's light nodes only).
template <typename MT>
__device__ inline void build_masks(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                   const int32_t* __restrict__ enodes,
                                   const uint32_t* __restrict__ voff,
                                   const int32_t* __restrict__ vedges,
                                   const int32_t* __restrict__ rank,
                                   const int8_t* __restrict__ hub = nullptr) {
  for (int x = threadIdx.x; x < E; x += blockDim.x) M[x] = (MT)0;
  __syncthreads();
  const uint32_t s = eoff[a], t = eoff[a + 1];
  for (uint32_t i = s; i < t; ++i) {
    const int32_t v = enodes[i];
    if (hub && hub[v] >= 0) continue;  // uniform across the workgroup
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

namespace hm {

// Index of the lowest set bit of a non-zero mask.
__device__ inline int low_bit(uint32_t m) { return __builtin_ctz(m); }
__device__ inline int low_bit(mask64_t m) { return __builtin_ctzll(m); }

// First position p in [lo, hi) with keys[p] >= t (keys ascending on [lo, hi)), else hi.
__device__ inline uint32_t lower_bound_key(const int32_t* __restrict__ keys, uint32_t lo, uint32_t hi,
                                           int32_t t) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The part of a hyperedge's V2H rows that falls in one key tile [k0, k1): for node a_i (i < sa),
// positions lo[i] .. lo[i] + (pre[i + 1] - pre[i]) of its row, whose keys (hyperedge ids or ranks,
// ascending within the row) lie in the tile. pre[sa] is the total.
struct TileRanges {
  uint32_t lo[HM_MAX_EDGE_SIZE];
  uint32_t pre[HM_MAX_EDGE_SIZE + 1];
};

// Fills tr for the nodes anodes[0 .. sa) (sa <= 64, handled by wave 0) and returns the total, the
// same on every lane. Starts and ends with a workgroup barrier, so it may follow any use of tr.
__device__ inline uint32_t tile_ranges(TileRanges& tr, int sa, const int32_t* __restrict__ anodes,
                                       const uint32_t* __restrict__ voff,
                                       const int32_t* __restrict__ keys, int32_t k0, int32_t k1) {
  __syncthreads();
  if (threadIdx.x < kWave) {
    const int i = (int)threadIdx.x;
    uint32_t n = 0;
    if (i < sa && k0 < k1) {
      const int32_t v = anodes[i];
      const uint32_t s = voff[v], t = voff[v + 1];
      const uint32_t lo = lower_bound_key(keys, s, t, k0);
      n = lower_bound_key(keys, lo, t, k1) - lo;
      tr.lo[i] = lo;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(n, o);
      if (i >= o) n += y;
    }
    if (i < sa) tr.pre[i + 1] = n;
    if (i == 0) tr.pre[0] = 0;
  }
  __syncthreads();
  return tr.pre[sa];
}

// Calls f(i, key) for every entry of the tile ranges, spread over all lanes of the workgroup.
template <typename F>
__device__ inline void for_tile_entries(const TileRanges& tr, int sa, uint32_t total,
                                        const int32_t* __restrict__ keys, F&& f) {
  for (uint32_t q = threadIdx.x; q < total; q += blockDim.x) {
    int lo = 0, hi = sa;  // largest i with pre[i] <= q
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tr.pre[mid] <= q) lo = mid;
      else hi = mid;
    }
    f(lo, keys[tr.lo[lo] + (q - tr.pre[lo])]);
  }
}

// Mask of a's nodes (s_an[0 .. sa), ascending) contained in hyperedge b.
template <typename MT>
__device__ inline MT pair_mask(const int32_t* s_an, int sa, const uint32_t* __restrict__ eoff,
                               const int32_t* __restrict__ enodes, int32_t b) {
  MT m = (MT)0;
  for (uint32_t q = eoff[b], qe = eoff[b + 1]; q < qe; ++q) {
    const int32_t v = enodes[q];
    int lo = 0, hi = sa;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_an[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    if (lo < sa && s_an[lo] == v) m |= (MT)1 << lo;
  }
  return m;
}

}  // namespace hm
