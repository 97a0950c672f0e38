// Dynamic hypergraphs: device-side batch insert/delete on the flat sorted arrays and the exact
// incremental (delta) h-motif count of a batch.
//
// Replaces the reference's CBST update path (dynamicHyperGraph/structure/operations.cu:82-453:
// fillCBST / unfillCBST / deleteCBST / insertCBST, kernels in dynamicHyperGraph/kernel/) and its
// delta count computeMotifCountsDelta (dynamicHyperGraph/include/motif_update.hpp:13-19,
// dynamicHyperGraph/src/HMotifCountUpdate.cu:99-301), whose anchor rule adds or subtracts every
// closed triangle anchored at a flagged edge and so is not the difference of two counts.
//
// Delta rule. For a set S of changed hyperedges in a graph G, the count of the triples of G that
// contain at least one member of S is split over S: a triple T is charged to min(T ∩ S) (by id).
// Deleted hyperedges are charged on the graph before the batch (subtracted), inserted ones on the
// graph after it (added), so  counts(after) = counts(before) - L(D, before) + L(A, after).
// L(S, G) is computed per changed hyperedge f (one workgroup) from the LDS mask table of f
// (M[x] = bitmask of f's nodes contained in x) and full projection rows N(b) of G:
//   * pairs {b, c} of neighbours of f: counted as if open at f from the neighbour classes
//     (C_fb, [b ⊄ f]), then every b < c with b ∩ c ≠ ∅ (c in N(b) ∩ N(f)) is re-booked from the
//     open bin to its closed bin;
//   * triples open at a neighbour b with f at one end: c in N(b) \ N(f), c ≠ f.
// Neighbours that are in S with a smaller id than f are skipped, which is the charging rule.
//
// Store update. Deleted ids keep an empty row (an empty hyperedge overlaps nothing, so it is in no
// triple); inserted hyperedges get fresh ids E, E+1, ... . H2V is rebuilt by a scan of the new
// sizes and a row copy; V2H by a stable radix sort of the (node, hyperedge) incidences, which
// keeps every node's hyperedge list ascending.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_tables.hpp"

namespace hm {

__constant__ int8_t c_delta_bin[128] = HM_STATIC_BIN_TABLE;

// Accumulator of a delta count (40 x uint64): bins 0..29 hold the signed delta (two's complement),
// slot 30 the triples charged to deleted hyperedges (before the batch), slot 31 the triples
// charged to inserted ones (after it), slot 32 the row entries the delta kernel scanned.
constexpr int kDynAccRemoved = 30;
constexpr int kDynAccAdded = 31;
constexpr int kDynAccEntries = 32;  // projection-row entries scanned by k_delta
constexpr int kDynAccCount = 40;

// projected degree d_a (all neighbours, by id)
template <typename MT>
__device__ inline void deg_body(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                const int32_t* __restrict__ enodes,
                                const uint32_t* __restrict__ voff,
                                const int32_t* __restrict__ vedges,
                                unsigned long long* __restrict__ deg) {
  __shared__ uint32_t red[kWavesPerBlock];
  if (eoff[a + 1] == eoff[a]) {  // empty (deleted) hyperedge
    if (threadIdx.x == 0) deg[a] = 0;
    return;
  }
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, nullptr);
  uint32_t d = 0;
  for (int32_t x = threadIdx.x; x < E; x += blockDim.x) d += M[x] != (MT)0;
  d = wave_sum(d);
  if (lane_id() == 0) red[wave_id()] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) t += red[q];
    deg[a] = t;
  }
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_deg(int32_t E, const uint32_t* __restrict__ eoff,
                                                const int32_t* __restrict__ enodes,
                                                const uint32_t* __restrict__ voff,
                                                const int32_t* __restrict__ vedges,
                                                unsigned long long* __restrict__ deg) {
  extern __shared__ __align__(16) unsigned char smem[];
  deg_body<MT>(reinterpret_cast<MT*>(smem), E, (int32_t)blockIdx.x, eoff, enodes, voff, vedges, deg);
}

// full row of a by id: (x << 7 | C_ax) for every x != a with C_ax > 0, ascending
template <typename MT>
__device__ inline void full_fill_body(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                      const int32_t* __restrict__ enodes,
                                      const uint32_t* __restrict__ voff,
                                      const int32_t* __restrict__ vedges,
                                      const unsigned long long* __restrict__ roff,
                                      uint32_t* __restrict__ rows) {
  __shared__ uint32_t wcnt[kWavesPerBlock];
  if (eoff[a + 1] == eoff[a]) return;
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, nullptr);
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t* out = rows + roff[a];
  uint32_t running = 0;
  for (int32_t x0 = 0; x0 < E; x0 += blockDim.x) {
    const int32_t x = x0 + (int32_t)threadIdx.x;
    const MT m = x < E ? M[x] : (MT)0;
    const unsigned long long bal = __ballot(m != (MT)0);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) {
      woff += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    if (m) out[running + woff + (uint32_t)__popcll(bal & below)] =
        ((uint32_t)x << kCntBits) | (uint32_t)popc(m);
    running += tot;
    __syncthreads();
  }
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_full_fill(int32_t E, const uint32_t* __restrict__ eoff,
                                                      const int32_t* __restrict__ enodes,
                                                      const uint32_t* __restrict__ voff,
                                                      const int32_t* __restrict__ vedges,
                                                      const unsigned long long* __restrict__ roff,
                                                      uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  full_fill_body<MT>(reinterpret_cast<MT*>(smem), E, (int32_t)blockIdx.x, eoff, enodes, voff, vedges,
                     roff, rows);
}

// L(S, G) for the changed hyperedges in `changed` (flag[x] = 1 iff x in S); adds sign * L to acc.
// Each changed hyperedge f is shared by `parts` workgroups of NW waves (workgroup (f, p) takes the
// neighbours b of f with index = p mod parts), so a batch of a few hundred hyperedges still fills
// the GPU; every workgroup builds f's mask table, part 0 also counts the pairs open at f.
// Histogram: per wave, kDeltaCopies interleaved int32 copies updated with per-lane LDS atomics.
constexpr int kDeltaWaves = 16;
constexpr int kDeltaCopies = 8;
// Row loads in flight per lane: many when the LDS table already leaves one workgroup per CU
// (the registers they take cost no occupancy), few when several workgroups fit per CU.
constexpr int kDeltaLoadsWide = 16;
constexpr int kDeltaLoadsNarrow = 2;
// The narrow form serves graphs whose table leaves room for two 16-wave workgroups per CU: at
// most 64 VGPRs (8 waves per SIMD) keeps both resident (at 85 VGPRs only one fits).
constexpr int kDeltaWavesPerEu(int loads) { return loads <= kDeltaLoadsNarrow ? 8 : 1; }
constexpr size_t kDeltaOneGroupLds = 80 * 1024;  // above this, one workgroup per CU (160 KB LDS)
constexpr int kDeltaTiledLoads = 4;  // row chunks per step of the tiled delta
// Slots of a closed pair (b, c) of a changed hyperedge f (Venn id v over f, b, c in that order:
// bits 0-2 private regions, 3-5 pairwise-only regions fb, bc, fc, 6 the triple region). Its
// as-if-open bin at f, 20 + A + 2 ([b ⊄ f] + [c ⊄ f]), is fixed by v ([b ⊄ f] is bit 1 or 4,
// [c ⊄ f] bit 2 or 4) and A = [|f| > C_fb + C_fc], which is bit 0 without a triple region.
// Slot = (low << 1 | A) + 16 mid, low = bits 0-2, mid = bits 3-5: 0..127 with a triple region,
// 144 + (low << 1 | A) without (mid = 7). Slots 128..143 count the triples (f, b, c) open at b
// (c ∩ f = ∅) by F | F << 1 | X << 2 | Y << 3, bin 20 + X + 2 (F + Y): F = [f ⊄ b],
// X = [b has a node in neither f nor c], Y = [c ⊄ b].
constexpr int kDSlots = 160;
constexpr int kDSlotOpenB = 128;
constexpr int kDSlotNoTriple = 144;
constexpr int kDSlotCopies = 32;  // copies per slot shared by the workgroup (copy = lane & 31)

// Venn id and as-if-open bin (at f) of closed-pair slot p (p < 128 or p >= kDSlotNoTriple).
__device__ inline void dslot_bins(int p, int* v, int* as_open) {
  const int q = p < 128 ? p : p - kDSlotNoTriple;
  const int a = q & 1, low = (q >> 1) & 7;
  const int vv = p < 128 ? (low | (p >> 4) << 3 | 64) : (low | 56);
  const int bf = ((vv >> 1) & 1) | ((vv >> 4) & 1), cf = ((vv >> 2) & 1) | ((vv >> 4) & 1);
  *v = vv;
  *as_open = kOpenBin0 + a + 2 * (bf + cf);
}
// Byte offset of the counter (slot, copy) of a row entry (b, c) of neighbour b of f, for 32
// copies per slot (copy fastest) and base = 4 copy. Each slot bit is [P > Q] over counts 0..64,
// bit 7 of the byte 127 + P - Q; word A holds f only [ffree + g > C_fc], b only
// [bfree + g > C_bc], c only [s_c + g > C_fc + C_bc] and A [ffree > C_fc], word B fb only
// [C_fb > g], bc only [C_bc > g], fc only [C_fc > g] and a constant 128 (ffree = |f| - C_fb,
// bfree = |b| - C_fb, g = |f ∩ b ∩ c|). KA holds b's constants (127 + ffree, 127 + bfree, 127,
// 127 + ffree) and KB (127 + C_fb, 127, 127, 128). Byte dot products with the slot bit weights
// gather the bits (times 128): with g = 0 the slot is kDSlotNoTriple + (low << 1 | A) for c
// overlapping f and kDSlotOpenB + (F | F << 1 | X << 2 | Y << 3) for c not (its f-only and A bits
// are then F, its b-only bit X, its c-only bit Y), told apart by the fc-only bit.
__device__ inline uint32_t dslot_offset(uint32_t KA, uint32_t KB, uint32_t cfc, uint32_t sc,
                                        uint32_t cbc, uint32_t g, uint32_t base) {
  const uint32_t gm = g * 0x10101u;
  const uint32_t c24 = __builtin_amdgcn_perm(0u, cfc, 0x000c0c0cu);  // C_fc << 24
  const uint32_t A = (KA + (sc << 16) + gm - cfc * 0x10001u - c24 - cbc * 0x10100u) & 0x80808080u;
  const uint32_t B = (KB + (cbc << 8) + (cfc << 16) - gm) & 0x80808080u;
  const uint32_t sA = __builtin_amdgcn_udot4(A, 2u | 4u << 8 | 8u << 16 | 1u << 24, base, false);
  return g ? __builtin_amdgcn_udot4(B, 16u | 32u << 8 | 64u << 16, sA, false)
           : __builtin_amdgcn_udot4(B, 16u << 16 | 128u << 24, sA, false);
}

constexpr size_t kDeltaTiledStaticLds = 40 * 1024;  // k_delta_tiled: histograms, classes, tables
constexpr size_t kDeltaStaticLds = 24 * 1024;       // k_delta: slot counters, bins, classes, tables

template <typename MT, int NW, int kDeltaLoads>
__device__ inline void delta_body(
    MT* M, int32_t E, int32_t f, uint32_t part, int32_t parts, const uint32_t* __restrict__ eoff,
    const int32_t* __restrict__ enodes, const uint8_t* __restrict__ esize,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const unsigned long long* __restrict__ roff, const unsigned long long* __restrict__ rlen,
    const uint32_t* __restrict__ rows, const uint8_t* __restrict__ flag, long long sign,
    unsigned long long* __restrict__ acc) {
  // size of x (at most 64) in bits 0..6, bit 7 set for x = f and for an x of S before f (their
  // triples with f belong to an earlier changed hyperedge): one LDS read per row entry
  uint8_t* szf = reinterpret_cast<uint8_t*>(M + E);
  __shared__ int32_t h[NW][32];                       // open bins, then every bin
  __shared__ uint32_t hs[kDSlots * kDSlotCopies];  // closed pairs by slot
  __shared__ uint32_t cls[2 * HM_MAX_EDGE_SIZE];
  __shared__ unsigned long long red[6][NW];
  __shared__ int8_t tbl[128];
  __shared__ uint32_t s_next;
  __shared__ unsigned long long s_scan;
  if (eoff[f + 1] == eoff[f]) return;
  for (int k = threadIdx.x; k < NW * 32; k += blockDim.x) (&h[0][0])[k] = 0;
  for (int k = threadIdx.x; k < kDSlots * kDSlotCopies; k += blockDim.x) hs[k] = 0u;
  for (int k = threadIdx.x; k < 2 * HM_MAX_EDGE_SIZE; k += blockDim.x) cls[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_delta_bin[threadIdx.x];
  if (threadIdx.x == 0) {
    s_next = 0;
    s_scan = 0;
  }
  for (int32_t x = threadIdx.x; x < E; x += blockDim.x)
    szf[x] = (uint8_t)(esize[x] | ((flag[x] && x < f) || x == f ? 0x80 : 0));
  build_masks<MT>(M, E, f, eoff, enodes, voff, vedges, nullptr);
  const int sf = (int)(eoff[f + 1] - eoff[f]);
  const int lane = lane_id(), w = wave_id();
  if (part == 0) {
    // classes of the allowed neighbours, then every pair of them as if open at f (64-bit, straight
    // to the global accumulator: C(d_f, 2) can exceed the int32 LDS counters)
    for (int32_t x = threadIdx.x; x < E; x += blockDim.x) {
      const MT m = M[x];
      const int zx = szf[x];
      if (m && !(zx & 0x80)) {
        const int C = popc(m);
        atomicAdd(&cls[2 * (C - 1) + ((zx & 0x7F) > C ? 1 : 0)], 1u);
      }
    }
    __syncthreads();
    const int K = 2 * sf;
    unsigned long long o[6] = {0, 0, 0, 0, 0, 0};
    for (int p = threadIdx.x; p < K * K; p += blockDim.x) {
      const int k1 = p / K, k2 = p - k1 * K;
      if (k1 > k2) continue;
      const unsigned long long n1 = cls[k1], n2 = cls[k2];
      const unsigned long long pairs = (k1 == k2) ? n1 * (n1 - (n1 > 0)) / 2 : n1 * n2;
      const int b6 = (int)((k1 >> 1) + (k2 >> 1) + 2 < sf) + 2 * ((k1 & 1) + (k2 & 1));
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] += b6 == q ? pairs : 0;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const unsigned long long v = wave_sum(o[q]);
      if (lane == 0) red[q][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < kWave) {  // wave 0: one atomic per bin, the touched total summed first
      unsigned long long t = 0;
      if (threadIdx.x < 6)
        for (int q = 0; q < NW; ++q) t += red[threadIdx.x][q];
      if (t) atomicAdd(&acc[kOpenBin0 + threadIdx.x], (unsigned long long)((long long)t * sign));
      const unsigned long long tot = wave_sum(t);
      if (threadIdx.x == 0 && tot) atomicAdd(&acc[sign < 0 ? kDynAccRemoved : kDynAccAdded], tot);
    }
  }
  // neighbours b of f (this part's share): closed pairs (b, c), b < c, re-booked, and triples
  // open at b (c overlaps b, not f); each is one LDS atomic into its slot (dslot_offset), which
  // fixes the closed bin (+1) and the as-if-open bin at f (-1) of a closed pair, or the open bin.
  unsigned char* hwb = reinterpret_cast<unsigned char*>(hs);
  const uint32_t cbase = 4u * (uint32_t)(lane & (kDSlotCopies - 1));
  const uint32_t* rowF = rows + roff[f];
  const uint32_t nF = (uint32_t)rlen[f];
  for (;;) {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(&s_next, 1u);
    k = __shfl(k, 0);
    const uint32_t ib = part + k * (uint32_t)parts;
    if (ib >= nF) break;
    const int32_t b = (int32_t)(rowF[ib] >> kCntBits);
    const uint32_t zb = szf[b];
    if (zb & 0x80u) continue;
    const MT mb = M[b];
    const uint32_t cfb = (uint32_t)popc(mb);
    const uint32_t ffree = (uint32_t)sf - cfb, bfree = zb - cfb;
    const uint32_t KA = (127u + ffree) | (127u + bfree) << 8 | 127u << 16 | (127u + ffree) << 24;
    const uint32_t KB = (127u + cfb) | 127u << 8 | 127u << 16 | 128u << 24;
    const uint32_t* rowB = rows + roff[b];
    const uint32_t nB = (uint32_t)rlen[b];
    if (lane == 0) atomicAdd(&s_scan, (unsigned long long)nB);
    auto visit = [&](uint32_t e) {
      const uint32_t c = e >> kCntBits;
      const uint32_t zc = szf[c];
      const MT mc = M[c];
      if ((zc & 0x80u) || (mc && c < (uint32_t)b)) return;
      atomicAdd(reinterpret_cast<uint32_t*>(hwb + dslot_offset(KA, KB, (uint32_t)popc(mc), zc, e & kCntMask,
                                                               (uint32_t)popc((MT)(mb & mc)), cbase)),
                1u);
    };
    // kDeltaLoads row loads in flight per lane before any is used (lane l reads l, l + 64, ...),
    // through a buffer descriptor: loads that wait only on the vector-memory counter (a generic
    // load also holds the LDS counter, so the first LDS read would wait for all of them)
    const __amdgpu_buffer_rsrc_t rs = row_rsrc(rowB, nB);
    uint32_t j = lane;
    for (; j + (kDeltaLoads - 1) * kWave < nB; j += kDeltaLoads * kWave) {
      uint32_t e[kDeltaLoads];
#pragma unroll
      for (int u = 0; u < kDeltaLoads; ++u) e[u] = row_load(rs, j + u * kWave);
#pragma unroll
      for (int u = 0; u < kDeltaLoads; ++u) visit(e[u]);
    }
    for (; j < nB; j += kWave) visit(row_load(rs, j));
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kDSlots; p += blockDim.x) {  // slots into their bins
    uint32_t n = 0;
    for (int c = 0; c < kDSlotCopies; ++c) n += hs[p * kDSlotCopies + c];
    if (!n) continue;
    if (p >= kDSlotOpenB && p < kDSlotNoTriple) {
      const int q = p - kDSlotOpenB;  // F | F << 1 | X << 2 | Y << 3
      atomicAdd(&h[0][kOpenBin0 + ((q >> 2) & 1) + 2 * ((q & 1) + ((q >> 3) & 1))], (int32_t)n);
    } else {
      int v, as_open;
      dslot_bins(p, &v, &as_open);
      atomicAdd(&h[0][tbl[v]], (int32_t)n);
      atomicSub(&h[0][as_open], (int32_t)n);
    }
  }
  __syncthreads();
  if (threadIdx.x < kWave) {  // wave 0: one atomic per bin; the touched total (two's complement,
                              // a bin's net count may be negative) summed across the wave first
    long long s = 0;
    if (threadIdx.x < kStaticBins)
      for (int q = 0; q < NW; ++q) s += h[q][threadIdx.x];
    if (s) atomicAdd(&acc[threadIdx.x], (unsigned long long)(s * sign));
    const unsigned long long tot = wave_sum((unsigned long long)s);
    if (threadIdx.x == 0 && tot) atomicAdd(&acc[sign < 0 ? kDynAccRemoved : kDynAccAdded], tot);
  }
  if (threadIdx.x == 0 && s_scan) atomicAdd(&acc[kDynAccEntries], s_scan);
}

template <typename MT, int NW, int kDeltaLoads>
__global__ __launch_bounds__(NW * kWave) __attribute__((amdgpu_waves_per_eu(kDeltaWavesPerEu(kDeltaLoads)))) void k_delta(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, const unsigned long long* __restrict__ roff,
    const unsigned long long* __restrict__ rlen, const uint32_t* __restrict__ rows,
    const int32_t* __restrict__ changed, const uint8_t* __restrict__ flag, int32_t parts,
    long long sign, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  delta_body<MT, NW, kDeltaLoads>(reinterpret_cast<MT*>(smem), E, changed[blockIdx.x / parts],
                                  blockIdx.x % parts, parts, eoff, enodes, esize, voff, vedges,
                                  roff, rlen, rows, flag, sign, acc);
}

// ---- tiled delta: hyperedge ids beyond one LDS table ------------------------------------------
// The id space is cut into tiles of T ids (the walk of hm_static's tiled kernels: the part of each
// node's V2H row inside the tile, found by binary search, and visit-once ownership by the lane of
// the lowest node). Full rows come out grouped by tile, so the entries of a row inside a tile are
// contiguous and a read position per (f, b) pair walks them tile after tile.

// Projected degree d_a over id tiles.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_deg_tiled(int32_t E, int32_t T, const uint32_t* __restrict__ eoff,
                                                      const int32_t* __restrict__ enodes,
                                                      const uint32_t* __restrict__ voff,
                                                      const int32_t* __restrict__ vedges,
                                                      unsigned long long* __restrict__ deg) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ TileRanges tr;
  __shared__ uint32_t red[kWavesPerBlock];
  const int32_t a = blockIdx.x;
  const uint32_t s = eoff[a];
  const int sa = (int)(eoff[a + 1] - s);
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  uint32_t d = 0;
  for (int32_t t0 = 0; t0 < E; t0 += T) {
    const int32_t t1 = min(E, t0 + T);
    const uint32_t total = tile_ranges(tr, sa, enodes + s, voff, vedges, t0, t1);
    if (total == 0) continue;
    for_tile_entries(tr, sa, total, vedges, [&](int i, int32_t x) {
      if (x != a) atomicOr(&M[x - t0], (MT)1 << i);
    });
    __syncthreads();
    for_tile_entries(tr, sa, total, vedges, [&](int i, int32_t x) {
      if (x == a) return;
      const MT m = M[x - t0];
      if (m && low_bit(m) == i) {
        ++d;
        M[x - t0] = (MT)0;
      }
    });
  }
  d = wave_sum(d);
  if (lane_id() == 0) red[wave_id()] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) t += red[q];
    deg[a] = t;
  }
}

// Full row of a, (x << 7 | C_ax) grouped by id tile in ascending tile order.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_full_fill_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const unsigned long long* __restrict__ roff, uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ TileRanges tr;
  __shared__ uint32_t s_cnt;
  const int32_t a = blockIdx.x;
  const uint32_t s = eoff[a];
  const int sa = (int)(eoff[a + 1] - s);
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  if (threadIdx.x == 0) s_cnt = 0;
  uint32_t* out = rows + roff[a];
  for (int32_t t0 = 0; t0 < E; t0 += T) {
    const int32_t t1 = min(E, t0 + T);
    const uint32_t total = tile_ranges(tr, sa, enodes + s, voff, vedges, t0, t1);
    if (total == 0) continue;
    for_tile_entries(tr, sa, total, vedges, [&](int i, int32_t x) {
      if (x != a) atomicOr(&M[x - t0], (MT)1 << i);
    });
    __syncthreads();
    for_tile_entries(tr, sa, total, vedges, [&](int i, int32_t x) {
      if (x == a) return;
      const MT m = M[x - t0];
      if (m && low_bit(m) == i) {
        out[atomicAdd(&s_cnt, 1u)] = ((uint32_t)x << kCntBits) | (uint32_t)popc(m);
        M[x - t0] = (MT)0;
      }
    });
  }
}

// rlen of every changed hyperedge (state sizes of k_delta_tiled).
__global__ void k_gather_len(const int32_t* __restrict__ changed, int32_t n,
                             const unsigned long long* __restrict__ rlen,
                             unsigned long long* __restrict__ out) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = rlen[changed[q]];
  else if (q == n) out[n] = 0;
}

// L(S, G) over id tiles: the k_delta decomposition with f's mask table built one tile at a time.
// Per neighbour b of f (this part's share) the state at soff[q] + ib keeps f's mask in b and the
// read position in b's tile-grouped row. Flags and sizes are read from HBM (they do not fit LDS).
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_delta_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, const unsigned long long* __restrict__ roff,
    const unsigned long long* __restrict__ rlen, const uint32_t* __restrict__ rows,
    const int32_t* __restrict__ changed, const uint8_t* __restrict__ flag, int32_t parts,
    long long sign, const unsigned long long* __restrict__ soff, uint32_t* __restrict__ cur,
    MT* __restrict__ pm, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ int32_t h[NW][32][kDeltaCopies];
  __shared__ uint32_t hs[kDSlots * kDSlotCopies];  // closed pairs by slot, as in k_delta
  __shared__ uint32_t cls[2 * HM_MAX_EDGE_SIZE];
  __shared__ unsigned long long red[6][NW];
  __shared__ int8_t tbl[128];
  __shared__ int32_t s_an[HM_MAX_EDGE_SIZE];
  __shared__ TileRanges tr;
  __shared__ unsigned long long s_scan;
  const int32_t q = blockIdx.x / parts;
  const int32_t f = changed[q];
  const uint32_t part = blockIdx.x % parts;
  const uint32_t fs = eoff[f];
  const int sf = (int)(eoff[f + 1] - fs);
  if (sf == 0) return;
  for (int k = threadIdx.x; k < NW * 32 * kDeltaCopies; k += blockDim.x) (&h[0][0][0])[k] = 0;
  for (int k = threadIdx.x; k < kDSlots * kDSlotCopies; k += blockDim.x) hs[k] = 0u;
  for (int k = threadIdx.x; k < 2 * HM_MAX_EDGE_SIZE; k += blockDim.x) cls[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_delta_bin[threadIdx.x];
  if ((int)threadIdx.x < sf) s_an[threadIdx.x] = enodes[fs + threadIdx.x];
  if (threadIdx.x == 0) s_scan = 0;
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  const int lane = lane_id(), w = wave_id();
  int32_t* hc = &h[w][0][lane & (kDeltaCopies - 1)];
  uint32_t* hsw = hs + (lane & (kDSlotCopies - 1));
  const uint32_t* rowF = rows + roff[f];
  const uint32_t nF = (uint32_t)rlen[f];
  const unsigned long long st = soff[q];
  for (int32_t t0 = 0; t0 < E; t0 += T) {
    const int32_t t1 = min(E, t0 + T);
    const uint32_t total = tile_ranges(tr, sf, enodes + fs, voff, vedges, t0, t1);
    for_tile_entries(tr, sf, total, vedges, [&](int i, int32_t x) {
      if (x != f) atomicOr(&M[x - t0], (MT)1 << i);
    });
    __syncthreads();
    if (part == 0)  // classes of the allowed neighbours inside this tile (each counted once)
      for_tile_entries(tr, sf, total, vedges, [&](int i, int32_t x) {
        if (x == f) return;
        const MT m = M[x - t0];
        if (m && low_bit(m) == i && (!flag[x] || x > f)) {
          const int C = popc(m);
          atomicAdd(&cls[2 * (C - 1) + (esize[x] > C ? 1 : 0)], 1u);
        }
      });
    // neighbours b of f (this part's share): closed pairs (re-booked) and triples open at b
    for (uint32_t k = (uint32_t)w;; k += NW) {
      const uint32_t ib = part + k * (uint32_t)parts;
      if (ib >= nF) break;
      const int32_t b = (int32_t)(rowF[ib] >> kCntBits);
      if (flag[b] && b < f) continue;
      MT mb;
      uint32_t pos;
      if (t0 == 0) {
        mb = pair_mask<MT>(s_an, sf, eoff, enodes, b);
        pos = 0;
        if (lane == 0) pm[st + ib] = mb;
      } else {
        mb = pm[st + ib];
        pos = cur[st + ib];
      }
      const int cfb = popc(mb);
      const int sb = esize[b];
      const uint32_t* rowB = rows + roff[b];
      const uint32_t nB = (uint32_t)rlen[b];
      const uint32_t pos0 = pos;
      // kDeltaTiledLoads chunks of b's row per step; the row's entries inside the tile are a
      // prefix, so each chunk's valid entries follow the previous chunk's
      for (;;) {
        uint32_t e[kDeltaTiledLoads];
        bool v[kDeltaTiledLoads];
#pragma unroll
        for (int u = 0; u < kDeltaTiledLoads; ++u) {
          const uint32_t j = pos + (uint32_t)(u * kWave) + lane;
          e[u] = j < nB ? rowB[j] : 0xFFFFFFFFu;
          v[u] = j < nB && (int32_t)(e[u] >> kCntBits) < t1;
        }
        uint32_t nv = 0;
#pragma unroll
        for (int u = 0; u < kDeltaTiledLoads; ++u) {
          if (v[u]) {
            const int32_t c = (int32_t)(e[u] >> kCntBits);
            if (c != f && !(flag[c] && c < f)) {
              const int cbc = (int)(e[u] & kCntMask);
              const MT mc = M[c - t0];
              const int sc = esize[c];
              if (mc) {
                if (c > b) {
                  const int cfc = popc(mc);
                  const int g = popc((MT)(mb & mc));
                  const unsigned low =
                      nz(sf - cfb - cfc + g) | nz(sb - cfb - cbc + g) << 1 | nz(sc - cfc - cbc + g) << 2;
                  const unsigned mid = nz(cfb - g) | nz(cbc - g) << 1 | nz(cfc - g) << 2;
                  const unsigned a = (unsigned)max(min(sf - cfb - cfc, 1), 0);
                  atomicAdd(&hsw[((low << 1 | a) + (mid << 4) + ((1u - nz(g)) << 5)) * kDSlotCopies], 1u);
                }
              } else {
                atomicAdd(&hc[open_bin(sb, sf, sc, cfb, cbc) * kDeltaCopies], 1);
              }
            }
          }
          nv += (uint32_t)__popcll(__ballot(v[u]));
        }
        pos += nv;
        if (nv < (uint32_t)(kDeltaTiledLoads * kWave)) break;
      }
      if (lane == 0) {
        cur[st + ib] = pos;
        if (pos > pos0) atomicAdd(&s_scan, (unsigned long long)(pos - pos0));
      }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < kDSlots; p += blockDim.x) {  // this tile's slots into their bins
      uint32_t n = 0;
      for (int c = 0; c < kDSlotCopies; ++c) {
        n += hs[p * kDSlotCopies + c];
        hs[p * kDSlotCopies + c] = 0u;
      }
      if (n) {
        int v, as_open;
        dslot_bins(p, &v, &as_open);
        atomicAdd(&h[0][tbl[v]][0], (int32_t)n);
        atomicSub(&h[0][as_open][0], (int32_t)n);
      }
    }
    __syncthreads();
    if (threadIdx.x < kWave) {  // flush this tile's histogram (bounds the int32 counters); the
                                // touched total summed across wave 0 first (one atomic, not 30)
      long long sum = 0;
      if (threadIdx.x < kStaticBins)
        for (int w2 = 0; w2 < NW; ++w2)
          for (int c = 0; c < kDeltaCopies; ++c) {
            sum += h[w2][threadIdx.x][c];
            h[w2][threadIdx.x][c] = 0;
          }
      if (sum) atomicAdd(&acc[threadIdx.x], (unsigned long long)(sum * sign));
      const unsigned long long tot = wave_sum((unsigned long long)sum);
      if (threadIdx.x == 0 && tot) atomicAdd(&acc[sign < 0 ? kDynAccRemoved : kDynAccAdded], tot);
    }
    for_tile_entries(tr, sf, total, vedges, [&](int, int32_t x) { M[x - t0] = (MT)0; });
  }
  if (part == 0) {  // every pair of allowed neighbours as if open at f (64-bit)
    __syncthreads();
    const int K = 2 * sf;
    unsigned long long o[6] = {0, 0, 0, 0, 0, 0};
    for (int p = threadIdx.x; p < K * K; p += blockDim.x) {
      const int k1 = p / K, k2 = p - k1 * K;
      if (k1 > k2) continue;
      const unsigned long long n1 = cls[k1], n2 = cls[k2];
      const unsigned long long pairs = (k1 == k2) ? n1 * (n1 - (n1 > 0)) / 2 : n1 * n2;
      const int b6 = (int)((k1 >> 1) + (k2 >> 1) + 2 < sf) + 2 * ((k1 & 1) + (k2 & 1));
#pragma unroll
      for (int qq = 0; qq < 6; ++qq) o[qq] += b6 == qq ? pairs : 0;
    }
#pragma unroll
    for (int qq = 0; qq < 6; ++qq) {
      const unsigned long long v = wave_sum(o[qq]);
      if (lane == 0) red[qq][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      unsigned long long t = 0;
      for (int w2 = 0; w2 < NW; ++w2) t += red[threadIdx.x][w2];
      if (t) {
        atomicAdd(&acc[kOpenBin0 + threadIdx.x], (unsigned long long)((long long)t * sign));
        atomicAdd(&acc[sign < 0 ? kDynAccRemoved : kDynAccAdded], t);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_scan) atomicAdd(&acc[kDynAccEntries], s_scan);
}

// ---- store update kernels -------------------------------------------------------------------
__global__ void k_new_sizes(int32_t E, int32_t nins, const uint32_t* __restrict__ eoff,
                            const uint8_t* __restrict__ del, const uint32_t* __restrict__ ins_off,
                            unsigned long long* __restrict__ sizes) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < E) sizes[e] = del[e] ? 0 : (eoff[e + 1] - eoff[e]);
  else if (e < E + nins) sizes[e] = ins_off[e - E + 1] - ins_off[e - E];
  else if (e == E + nins) sizes[e] = 0;
}

__global__ void k_copy_rows(int32_t E, int32_t nins, const uint32_t* __restrict__ eoff,
                            const int32_t* __restrict__ enodes, const uint8_t* __restrict__ del,
                            const uint32_t* __restrict__ ins_off, const int32_t* __restrict__ ins_nodes,
                            const unsigned long long* __restrict__ noff, int32_t* __restrict__ nnodes,
                            uint32_t* __restrict__ neoff, uint8_t* __restrict__ nsize,
                            int32_t* __restrict__ inc_node, int32_t* __restrict__ inc_edge) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > E + nins) return;
  neoff[e] = (uint32_t)noff[e];
  if (e == E + nins) return;
  const int32_t* src;
  uint32_t n;
  if (e < E) {
    src = enodes + eoff[e];
    n = del[e] ? 0 : eoff[e + 1] - eoff[e];
  } else {
    src = ins_nodes + ins_off[e - E];
    n = ins_off[e - E + 1] - ins_off[e - E];
  }
  const unsigned long long o = noff[e];
  nsize[e] = (uint8_t)(n > 255 ? 255 : n);
  for (uint32_t k = 0; k < n; ++k) {
    nnodes[o + k] = src[k];
    inc_node[o + k] = src[k];
    inc_edge[o + k] = e;
  }
}

// V2H offsets from the incidences sorted by node: off[v] = the first sorted position whose node
// is >= v, for v in [0, V] (keys at or above V, the padding of a captured update, count as V).
// Thread i (0..n) writes the offsets of the nodes in (key[i-1], key[i]], with key[-1] = -1 and
// key[n] = V, so each of the V + 1 offsets is written exactly once and no counter is shared
// (per-incidence atomics serialised on hub nodes: 4.1 ms per batch at 3.9M incidences).
__global__ void k_offsets_from_sorted(int64_t n, const uint32_t* __restrict__ key, int32_t V,
                                      uint32_t* __restrict__ off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int64_t cur = i < n ? min((int64_t)key[i], (int64_t)V) : (int64_t)V;
  const int64_t prev = i > 0 ? min((int64_t)key[i - 1], (int64_t)V) : -1;
  for (int64_t v = prev + 1; v <= cur; ++v) off[v] = (uint32_t)i;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void k_set_flags(const int32_t* __restrict__ ids, int32_t n, uint8_t* __restrict__ flag) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flag[ids[k]] = 1;
}

__global__ void k_zero_acc(unsigned long long* p, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

// ---- captured update: the same steps with the batch's sizes read from the device header ------
// The grids span the plan's capacities; workgroups and threads past the batch's sizes leave at
// once, so one instantiated graph serves every batch that fits the plan.

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_deg_cap(const UpdHdr* __restrict__ hd,
                                                    const uint32_t* __restrict__ eoff,
                                                    const int32_t* __restrict__ enodes,
                                                    const uint32_t* __restrict__ voff,
                                                    const int32_t* __restrict__ vedges,
                                                    unsigned long long* __restrict__ deg) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t E = hd->E_new, a = (int32_t)blockIdx.x;
  if (a >= E) {  // ids past the graph (up to the capacity) scan as empty rows
    if (threadIdx.x == 0) deg[a] = 0;
    return;
  }
  deg_body<MT>(reinterpret_cast<MT*>(smem), E, a, eoff, enodes, voff, vedges, deg);
}

// rows after the batch over capacity: flag it, so the fill and the addition phase are skipped
__global__ void k_check_rows_cap(UpdHdr* __restrict__ hd, const unsigned long long* __restrict__ roff,
                                 unsigned long long Pcap) {
  if (threadIdx.x == 0) {
    const unsigned long long P = roff[hd->E_new];
    hd->P_new = P;
    if (P > Pcap) hd->overflow = 1;
  }
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_full_fill_cap(const UpdHdr* __restrict__ hd,
                                                          const uint32_t* __restrict__ eoff,
                                                          const int32_t* __restrict__ enodes,
                                                          const uint32_t* __restrict__ voff,
                                                          const int32_t* __restrict__ vedges,
                                                          const unsigned long long* __restrict__ roff,
                                                          uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t E = hd->E_new, a = (int32_t)blockIdx.x;
  if (hd->overflow || a >= E) return;
  full_fill_body<MT>(reinterpret_cast<MT*>(smem), E, a, eoff, enodes, voff, vedges, roff, rows);
}

// phase 0: L(deleted ids, graph before the batch), subtracted; phase 1: L(inserted ids, graph
// after it), added (skipped when the rows overflowed)
template <typename MT, int NW, int kDeltaLoads>
__global__ __launch_bounds__(NW * kWave) __attribute__((amdgpu_waves_per_eu(kDeltaWavesPerEu(kDeltaLoads)))) void k_delta_cap(
    const UpdHdr* __restrict__ hd, int32_t phase, const int32_t* __restrict__ dels,
    const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, const unsigned long long* __restrict__ roff,
    const unsigned long long* __restrict__ rlen, const uint32_t* __restrict__ rows,
    const uint8_t* __restrict__ flag, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (phase && hd->overflow) return;
  const int32_t n = phase ? hd->n_ins : hd->n_del;
  const int32_t parts = phase ? hd->parts_ins : hd->parts_del;
  const int32_t q = (int32_t)(blockIdx.x / (uint32_t)parts);
  if (q >= n) return;
  const int32_t f = phase ? hd->E_old + q : dels[q];
  delta_body<MT, NW, kDeltaLoads>(reinterpret_cast<MT*>(smem), phase ? hd->E_new : hd->E_old, f,
                                  blockIdx.x % (uint32_t)parts, parts, eoff, enodes, esize, voff,
                                  vedges, roff, rlen, rows, flag, phase ? 1ll : -1ll, acc);
}

__global__ void k_set_flags_cap(const UpdHdr* __restrict__ hd, int32_t phase,
                                const int32_t* __restrict__ dels, uint8_t* __restrict__ flag) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (phase ? hd->n_ins : hd->n_del)) flag[phase ? hd->E_old + k : dels[k]] = 1;
}

__global__ void k_new_sizes_cap(const UpdHdr* __restrict__ hd, int32_t Ecap,
                                const uint32_t* __restrict__ eoff, const uint8_t* __restrict__ del,
                                const uint32_t* __restrict__ ins_off,
                                unsigned long long* __restrict__ sizes) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > Ecap) return;
  const int32_t E = hd->E_old, nins = hd->n_ins;
  if (e < E) sizes[e] = del[e] ? 0 : (eoff[e + 1] - eoff[e]);
  else if (e < E + nins) sizes[e] = ins_off[e - E + 1] - ins_off[e - E];
  else sizes[e] = 0;
}

__global__ void k_copy_rows_cap(UpdHdr* __restrict__ hd, int32_t Ecap, const uint32_t* __restrict__ eoff,
                                const int32_t* __restrict__ enodes, const uint8_t* __restrict__ del,
                                const uint32_t* __restrict__ ins_off,
                                const int32_t* __restrict__ ins_nodes,
                                const unsigned long long* __restrict__ noff,
                                int32_t* __restrict__ nnodes, uint32_t* __restrict__ neoff,
                                uint8_t* __restrict__ nsize, int32_t* __restrict__ inc_node,
                                int32_t* __restrict__ inc_edge) {
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > Ecap) return;
  const int32_t E = hd->E_old, E2 = E + hd->n_ins;
  neoff[e] = (uint32_t)noff[e];
  if (e == E2) hd->I_new = noff[e];
  if (e >= E2) {
    if (e < Ecap) nsize[e] = 0;
    return;
  }
  const int32_t* src;
  uint32_t n;
  if (e < E) {
    src = enodes + eoff[e];
    n = del[e] ? 0 : eoff[e + 1] - eoff[e];
  } else {
    src = ins_nodes + ins_off[e - E];
    n = ins_off[e - E + 1] - ins_off[e - E];
  }
  const unsigned long long o = noff[e];
  nsize[e] = (uint8_t)(n > 255 ? 255 : n);
  for (uint32_t k = 0; k < n; ++k) {
    nnodes[o + k] = src[k];
    inc_node[o + k] = src[k];
    inc_edge[o + k] = e;
  }
}

// incidences past the batch's total sort last (key above every node id)
__global__ void k_pad_inc_cap(const UpdHdr* __restrict__ hd, int64_t Icap, int32_t* __restrict__ inc_node,
                              int32_t* __restrict__ inc_edge, int32_t pad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < Icap && k >= (int64_t)hd->I_new) {
    inc_node[k] = pad;
    inc_edge[k] = 0;
  }
}


__global__ void k_copy_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

__global__ void k_copy_bytes(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

static unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

static DynScratch& dyn_scratch(hm_graph* g) { return g->dyn; }

// Workgroups per changed hyperedge of a delta launch over n changed ids: enough to give every CU
// several.
static int32_t delta_parts(int32_t n) {
  return n <= 0 ? 1 : std::max<int32_t>(1, std::min<int32_t>(64, (2048 + n - 1) / n));
}

// Full projection rows of the graph currently on the handle (T > 0: tiled kernels, rows grouped
// by id tile; T == 0: one LDS table, rows ascending).
template <typename MT>
static hm_status build_full_rows(hm_graph* g, DynScratch& d, int32_t T) {
  DeviceCsr& c = g->csr;
  const int32_t E = c.E;
  hipStream_t st = g->stream;
  const size_t lds = ((size_t)(T > 0 ? T : E) * sizeof(MT) + 15) & ~(size_t)15;
  HM_TRY(d.deg.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(d.roff.ensure(sizeof(unsigned long long) * (E + 1)));
  k_zero_acc<<<1, 1, 0, st>>>(d.deg.as<unsigned long long>() + E, 1);
  if (T > 0) {
    HM_TRY(allow_lds((const void*)k_deg_tiled<MT>, lds));
    k_deg_tiled<MT><<<E, kBlock, lds, st>>>(E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                            c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                            d.deg.as<unsigned long long>());
  } else {
    HM_TRY(allow_lds((const void*)k_deg<MT>, lds));
    k_deg<MT><<<E, kBlock, lds, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                      c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                      d.deg.as<unsigned long long>());
  }
  HM_HIP(hipGetLastError());
  size_t tb = 0;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d.deg.as<unsigned long long>(),
                                          d.roff.as<unsigned long long>(), E + 1, st));
  HM_TRY(d.tmp.ensure(tb));
  tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(d.tmp.p, tb, d.deg.as<unsigned long long>(),
                                          d.roff.as<unsigned long long>(), E + 1, st));
  unsigned long long P = 0;
  HM_HIP(hipMemcpyAsync(&P, d.roff.as<unsigned long long>() + E, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_TRY(d.rows.ensure(sizeof(uint32_t) * std::max<unsigned long long>(P, 1)));
  d.rows_P = P;
  if (T > 0) {
    HM_TRY(allow_lds((const void*)k_full_fill_tiled<MT>, lds));
    k_full_fill_tiled<MT><<<E, kBlock, lds, st>>>(E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                                  c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                                  d.roff.as<unsigned long long>(), d.rows.as<uint32_t>());
  } else {
    HM_TRY(allow_lds((const void*)k_full_fill<MT>, lds));
    k_full_fill<MT><<<E, kBlock, lds, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                            c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                            d.roff.as<unsigned long long>(), d.rows.as<uint32_t>());
  }
  HM_HIP(hipGetLastError());
  return HM_OK;
}

// acc (device, 40 x u64, two's complement) += sign * L(ids, current graph). Builds the full rows
// unless those of the current graph (same tile) are still held. evb >= 0: records the start, the
// end of the row build and the end of the delta kernel in g->ev[evb..evb+2] (read by
// add_phase_ms once the stream has drained); nothing waits here.
template <typename MT>
static hm_status local_count(hm_graph* g, DynScratch& d, const int32_t* ids, int32_t n,
                             long long sign, int32_t T, int evb) {
  if (n == 0) return HM_OK;
  DeviceCsr& c = g->csr;
  const int32_t E = c.E;
  hipStream_t st = g->stream;
  if (evb >= 0) HM_HIP(hipEventRecord(g->ev[evb], st));
  if (!d.rows_valid || d.rows_tile != T) {  // else: the rows of this graph are already built
    d.rows_valid = false;
    HM_TRY(build_full_rows<MT>(g, d, T));
    d.rows_valid = true;
    d.rows_tile = T;
  }
  if (evb >= 0) HM_HIP(hipEventRecord(g->ev[evb + 1], st));
  HM_TRY(d.flag.ensure((size_t)std::max(E, 1)));
  HM_TRY(d.ids.ensure(sizeof(int32_t) * n));
  HM_HIP(hipMemcpyAsync(d.ids.p, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  k_fill_u8<<<grid_for(E), kBlock, 0, st>>>(d.flag.as<uint8_t>(), E, 0);
  k_set_flags<<<grid_for(n), kBlock, 0, st>>>(d.ids.as<int32_t>(), n, d.flag.as<uint8_t>());
  const int32_t parts = delta_parts(n);
  if (T > 0) {
    // state per (changed f, neighbour b) at the exclusive scan of the changed rows' lengths
    HM_TRY(d.slen.ensure(sizeof(unsigned long long) * ((size_t)n + 1)));
    HM_TRY(d.soff.ensure(sizeof(unsigned long long) * ((size_t)n + 1)));
    k_gather_len<<<grid_for((int64_t)n + 1), kBlock, 0, st>>>(d.ids.as<int32_t>(), n, d.deg.as<unsigned long long>(),
                                                        d.slen.as<unsigned long long>());
    size_t tb = 0;
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d.slen.as<unsigned long long>(),
                                            d.soff.as<unsigned long long>(), n + 1, st));
    HM_TRY(d.tmp.ensure(tb));
    tb = d.tmp.bytes;
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(d.tmp.p, tb, d.slen.as<unsigned long long>(),
                                            d.soff.as<unsigned long long>(), n + 1, st));
    unsigned long long U = 0;
    std::vector<unsigned long long> lens((size_t)n);
    HM_HIP(hipMemcpyAsync(&U, d.soff.as<unsigned long long>() + n, 8, hipMemcpyDeviceToHost, st));
    HM_HIP(hipMemcpyAsync(lens.data(), d.slen.p, sizeof(unsigned long long) * (size_t)n, hipMemcpyDeviceToHost, st));
    HM_HIP(hipStreamSynchronize(st));
    // k_delta_tiled folds its counters into int32 once per tile: workgroup (f, part) takes at
    // most ceil(d_f / parts) neighbours b of f, each with at most T row entries inside a tile
    unsigned long long dmax = 0;
    for (unsigned long long x : lens) dmax = std::max(dmax, x);
    if ((dmax / (unsigned long long)parts + 1) * (unsigned long long)T >= (1ull << 31)) {
      set_error("delta: a changed hyperedge with %llu neighbours over %d parts x tile of %d ids "
                "exceeds the bound of the int32 per-tile counters (2^31)", dmax, parts, T);
      return HM_ERR_UNSUPPORTED;
    }
    HM_TRY(d.scur.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(U, 1)));
    HM_TRY(d.spm.ensure(sizeof(MT) * (size_t)std::max<unsigned long long>(U, 1)));
    const size_t lds = ((size_t)T * sizeof(MT) + 15) & ~(size_t)15;
    HM_TRY(allow_lds((const void*)k_delta_tiled<MT, kDeltaWaves>, lds));
    k_delta_tiled<MT, kDeltaWaves><<<(unsigned)n * (unsigned)parts, kDeltaWaves * kWave, lds, st>>>(
        E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
        c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), d.roff.as<unsigned long long>(),
        d.deg.as<unsigned long long>(), d.rows.as<uint32_t>(), d.ids.as<int32_t>(),
        d.flag.as<uint8_t>(), parts, sign, d.soff.as<unsigned long long>(), d.scur.as<uint32_t>(),
        d.spm.as<MT>(), d.acc.as<unsigned long long>());
  } else {
    const size_t lds = (((size_t)E * sizeof(MT) + (size_t)E) + 15) & ~(size_t)15;
    const bool one_group = lds + kDeltaStaticLds > kDeltaOneGroupLds;
    // most work first on graphs whose table leaves one workgroup per CU: the changed hyperedges
    // by projected degree, descending (a hub hyperedge's workgroups scan ~1000x the rows of a
    // small one's; launched first, they do not form the tail). Small graphs keep id order: there
    // the sort's launches cost more than the tail.
    const int32_t* changed = d.ids.as<int32_t>();
    if (one_group) {
      HM_TRY(d.slen.ensure(sizeof(unsigned long long) * ((size_t)n + 1)));
      HM_TRY(d.soff.ensure(sizeof(unsigned long long) * ((size_t)n + 1)));
      HM_TRY(d.ord.ensure(sizeof(int32_t) * (size_t)n));
      k_gather_len<<<grid_for((int64_t)n + 1), kBlock, 0, st>>>(d.ids.as<int32_t>(), n, d.deg.as<unsigned long long>(),
                                                                d.slen.as<unsigned long long>());
      size_t tb = 0;
      HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, d.slen.as<unsigned long long>(),
                                                          d.soff.as<unsigned long long>(), d.ids.as<int32_t>(),
                                                          d.ord.as<int32_t>(), n, 0, 64, st));
      HM_TRY(d.tmp.ensure(tb));
      tb = d.tmp.bytes;
      HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(d.tmp.p, tb, d.slen.as<unsigned long long>(),
                                                          d.soff.as<unsigned long long>(), d.ids.as<int32_t>(),
                                                          d.ord.as<int32_t>(), n, 0, 64, st));
      changed = d.ord.as<int32_t>();
    }
    auto fn = one_group ? k_delta<MT, kDeltaWaves, kDeltaLoadsWide> : k_delta<MT, kDeltaWaves, kDeltaLoadsNarrow>;
    HM_TRY(allow_lds((const void*)fn, lds));
    fn<<<(unsigned)n * (unsigned)parts, kDeltaWaves * kWave, lds, st>>>(
        E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
        c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), d.roff.as<unsigned long long>(),
        d.deg.as<unsigned long long>(), d.rows.as<uint32_t>(), changed,
        d.flag.as<uint8_t>(), parts, sign, d.acc.as<unsigned long long>());
  }
  HM_HIP(hipGetLastError());
  if (evb >= 0) HM_HIP(hipEventRecord(g->ev[evb + 2], st));
  return HM_OK;
}

// Adds the row-build and delta-kernel times of a local_count recorded at evb (stream drained).
static hm_status add_phase_ms(hm_graph* g, int evb) {
  float a = 0.f, b = 0.f;
  HM_HIP(hipEventElapsedTime(&a, g->ev[evb], g->ev[evb + 1]));
  HM_HIP(hipEventElapsedTime(&b, g->ev[evb + 1], g->ev[evb + 2]));
  g->ms_upd_rows += a;
  g->ms_upd_delta += b;
  return HM_OK;
}

static bool dense_fits(hm_graph* g, int32_t E, bool wide) {
  int lim = 65536;
  (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device);
  // k_delta: mask and size-and-flag byte per id, plus about 22 KB of static arrays (24 KB reserved)
  return (size_t)E * ((wide ? 8 : 4) + 1) + 16 + kDeltaStaticLds <= (size_t)lim;
}

static hm_status local_count_any(hm_graph* g, DynScratch& d, const int32_t* ids, int32_t n,
                                 long long sign, int evb) {
  if (g->csr.max_size > HM_MAX_EDGE_SIZE)
    return fail(HM_ERR_UNSUPPORTED, "delta count: hyperedge larger than 64 nodes");
  const bool wide = g->csr.max_size > 32;
  int32_t T = 0;  // one LDS table over all ids when it fits, else tiles as large as LDS allows
  if (g->tile > 0 || !dense_fits(g, g->csr.E, wide)) {
    int lim = 65536;
    (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device);
    const int32_t most = (int32_t)((((size_t)lim - kDeltaTiledStaticLds) / (wide ? 8 : 4)) & ~(size_t)63);
    T = g->tile > 0 ? g->tile : most;
    if (T > most) {
      set_error("delta count: a tile of %d ids exceeds the LDS limit", T);
      return HM_ERR_BAD_INPUT;
    }
  }
  return wide ? local_count<mask64_t>(g, d, ids, n, sign, T, evb)
              : local_count<uint32_t>(g, d, ids, n, sign, T, evb);
}

// Applies the batch to the device arrays (deleted rows emptied, inserted rows appended).
static hm_status apply_batch(hm_graph* g, DynScratch& d, const std::vector<uint8_t>& delmask,
                             const std::vector<uint32_t>& ins_off, const std::vector<int32_t>& ins_nodes,
                             int32_t nins, int32_t newV) {
  DeviceCsr& c = g->csr;
  const int32_t E = c.E, E2 = E + nins;
  hipStream_t st = g->stream;
  const int64_t Iins = ins_off.back();
  HM_TRY(d.flag.ensure((size_t)std::max(E, 1)));
  HM_HIP(hipMemcpyAsync(d.flag.p, delmask.data(), (size_t)E, hipMemcpyHostToDevice, st));
  HM_TRY(d.ins_off.ensure(sizeof(uint32_t) * ins_off.size()));
  HM_TRY(d.ins_nodes.ensure(sizeof(int32_t) * std::max<int64_t>(Iins, 1)));
  HM_HIP(hipMemcpyAsync(d.ins_off.p, ins_off.data(), sizeof(uint32_t) * ins_off.size(), hipMemcpyHostToDevice, st));
  if (Iins > 0)
    HM_HIP(hipMemcpyAsync(d.ins_nodes.p, ins_nodes.data(), sizeof(int32_t) * Iins, hipMemcpyHostToDevice, st));
  HM_TRY(d.sizes.ensure(sizeof(unsigned long long) * (E2 + 1)));
  HM_TRY(d.noff.ensure(sizeof(unsigned long long) * (E2 + 1)));
  k_new_sizes<<<grid_for(E2 + 1), kBlock, 0, st>>>(E, nins, c.eoff.as<uint32_t>(), d.flag.as<uint8_t>(),
                                                   d.ins_off.as<uint32_t>(), d.sizes.as<unsigned long long>());
  size_t tb = 0;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d.sizes.as<unsigned long long>(),
                                          d.noff.as<unsigned long long>(), E2 + 1, st));
  HM_TRY(d.tmp.ensure(tb));
  tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(d.tmp.p, tb, d.sizes.as<unsigned long long>(),
                                          d.noff.as<unsigned long long>(), E2 + 1, st));
  unsigned long long I2 = 0;
  HM_HIP(hipMemcpyAsync(&I2, d.noff.as<unsigned long long>() + E2, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  if (I2 >= (1ull << 32)) return fail(HM_ERR_OVER_BUDGET, "update: more than 2^32 incidences");
  const size_t I2b = (size_t)std::max<unsigned long long>(I2, 1);
  HM_TRY(d.nnodes.ensure(sizeof(int32_t) * I2b));
  HM_TRY(d.neoff.ensure(sizeof(uint32_t) * (E2 + 1)));
  HM_TRY(d.nsize.ensure((size_t)std::max(E2, 1)));
  HM_TRY(d.inc_node.ensure(sizeof(int32_t) * I2b));
  HM_TRY(d.inc_edge.ensure(sizeof(int32_t) * I2b));
  HM_TRY(d.srt_node.ensure(sizeof(int32_t) * I2b));
  HM_TRY(d.srt_edge.ensure(sizeof(int32_t) * I2b));
  HM_TRY(d.nvoff.ensure(sizeof(uint32_t) * ((size_t)newV + 1)));
  k_copy_rows<<<grid_for(E2 + 1), kBlock, 0, st>>>(
      E, nins, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), d.flag.as<uint8_t>(),
      d.ins_off.as<uint32_t>(), d.ins_nodes.as<int32_t>(), d.noff.as<unsigned long long>(),
      d.nnodes.as<int32_t>(), d.neoff.as<uint32_t>(), d.nsize.as<uint8_t>(),
      d.inc_node.as<int32_t>(), d.inc_edge.as<int32_t>());
  HM_HIP(hipGetLastError());
  // V2H: stable sort of incidences by node (hyperedge ids stay ascending within a node)
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) <= newV) ++bits;
  tb = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, d.inc_node.as<uint32_t>(), d.srt_node.as<uint32_t>(),
                                            d.inc_edge.as<int32_t>(), d.srt_edge.as<int32_t>(),
                                            (int)I2, 0, bits, st));
  HM_TRY(d.tmp.ensure(tb));
  tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(d.tmp.p, tb, d.inc_node.as<uint32_t>(), d.srt_node.as<uint32_t>(),
                                            d.inc_edge.as<int32_t>(), d.srt_edge.as<int32_t>(),
                                            (int)I2, 0, bits, st));
  k_offsets_from_sorted<<<grid_for((int64_t)I2 + 1), kBlock, 0, st>>>(
      (int64_t)I2, d.srt_node.as<uint32_t>(), newV, d.nvoff.as<uint32_t>());
  // swap the new arrays in
  std::swap(c.enodes, d.nnodes);
  std::swap(c.eoff, d.neoff);
  std::swap(c.esize, d.nsize);
  std::swap(c.vedges, d.srt_edge);
  std::swap(c.voff, d.nvoff);
  c.E = E2;
  c.V = newV;
  c.I = (int64_t)I2;
  // no wait here: the caller's host arrays stay alive until its stream has drained
  return HM_OK;
}

// ---- captured update (hm_graph_set_update_capture) -------------------------------------------
// One hipGraph holds the whole update with a delta: the removal delta over the rows held from the
// previous update, the store update, the projection rows of the graph after the batch and the
// addition delta. The batch's sizes travel in a device header; every buffer is sized to the plan's
// capacities, so the launch has no host wait. Rows after the batch that exceed the plan are
// caught on the device (k_check_rows_cap): the fill and the addition phase are skipped, and the
// host reruns the addition phase eagerly with larger buffers.

// Byte offsets of a plan's input block: header, deleted ids, insert offsets, inserted nodes.
struct InLayout {
  size_t dels, ins_off, ins_nodes, total;
};

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

static InLayout in_layout(const UpdPlan& p) {
  InLayout L;
  L.dels = align16(sizeof(UpdHdr));
  L.ins_off = align16(L.dels + sizeof(int32_t) * (size_t)p.dcap);
  L.ins_nodes = align16(L.ins_off + sizeof(uint32_t) * ((size_t)p.icap + 1));
  L.total = align16(L.ins_nodes + sizeof(int32_t) * (size_t)std::max<int64_t>(p.iicap, 1));
  return L;
}

constexpr size_t kUpdOutBytes = sizeof(unsigned long long) * kDynAccCount + sizeof(UpdHdr);

// Device addresses a captured update reads or writes; the graph is valid while they are unchanged.
static std::vector<const void*> plan_sig(hm_graph* g) {
  DeviceCsr& c = g->csr;
  DynScratch& d = g->dyn;
  return {c.eoff.p,    c.enodes.p,   c.esize.p,    c.voff.p,     c.vedges.p, d.deg.p,
          d.roff.p,    d.rows.p,     d.flag.p,     d.acc.p,      d.tmp.p,    d.sizes.p,
          d.noff.p,    d.nnodes.p,   d.neoff.p,    d.nsize.p,    d.inc_node.p, d.inc_edge.p,
          d.srt_node.p, d.srt_edge.p, d.nvoff.p,    g->uplan.in.p};
}

// Grows b to at least `need` bytes, keeping its contents.
static hm_status grow_keep(DevBuf& b, size_t need, hipStream_t st) {
  if (b.p && b.bytes >= need) return HM_OK;
  DevBuf nb;
  HM_TRY(nb.ensure(need));
  if (b.p && b.bytes) HM_HIP(hipMemcpyAsync(nb.p, b.p, b.bytes, hipMemcpyDeviceToDevice, st));
  HM_HIP(hipStreamSynchronize(st));
  b.release();
  b = nb;
  return HM_OK;
}

// Largest id count whose delta table (mask and size-and-flag byte per id) fits one workgroup.
static int32_t dense_max_ids(hm_graph* g, bool wide) {
  int lim = 65536;
  (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device);
  const long long room = (long long)lim - 16 - (long long)kDeltaStaticLds;
  return room <= 0 ? 0 : (int32_t)(room / ((wide ? 8 : 4) + 1));
}

// Largest delta grid over n <= ncap changed ids.
static unsigned delta_grid_cap(int32_t ncap) {
  long long best = 1;
  for (int32_t n = 1; n <= ncap; ++n) best = std::max<long long>(best, (long long)n * delta_parts(n));
  return (unsigned)best;
}

template <typename MT>
static size_t delta_lds(int32_t E) {
  return (((size_t)E * sizeof(MT) + (size_t)E) + 15) & ~(size_t)15;
}

template <typename MT>
static size_t table_lds(int32_t E) {
  return ((size_t)E * sizeof(MT) + 15) & ~(size_t)15;
}

template <typename MT>
static bool delta_one_group(int32_t Ecap) {
  return delta_lds<MT>(Ecap) + kDeltaStaticLds > kDeltaOneGroupLds;
}

// Enqueues the whole captured update on g->stream (under capture: nothing waits).
template <typename MT>
static hm_status enqueue_update_cap(hm_graph* g) {
  DeviceCsr& c = g->csr;
  DynScratch& d = g->dyn;
  UpdPlan& p = g->uplan;
  hipStream_t st = g->stream;
  const InLayout L = in_layout(p);
  unsigned char* in = p.in.as<unsigned char>();
  UpdHdr* hd = reinterpret_cast<UpdHdr*>(in);
  const int32_t* dels = reinterpret_cast<const int32_t*>(in + L.dels);
  const uint32_t* ins_off = reinterpret_cast<const uint32_t*>(in + L.ins_off);
  const int32_t* ins_nodes = reinterpret_cast<const int32_t*>(in + L.ins_nodes);
  unsigned long long* acc = d.acc.as<unsigned long long>();
  const int32_t Ecap = p.Ecap, Vcap = p.Vcap;
  const int64_t Icap = p.Icap;
  const size_t lds_d = delta_lds<MT>(Ecap), lds_t = table_lds<MT>(Ecap);
  auto kd = delta_one_group<MT>(Ecap) ? k_delta_cap<MT, kDeltaWaves, kDeltaLoadsWide>
                                      : k_delta_cap<MT, kDeltaWaves, kDeltaLoadsNarrow>;
  k_zero_acc<<<1, kDynAccCount, 0, st>>>(acc, kDynAccCount);
  // removal phase: L(deleted ids, graph before the batch) over the rows held from the last update
  k_fill_u8<<<grid_for(Ecap), kBlock, 0, st>>>(d.flag.as<uint8_t>(), Ecap, 0);
  k_set_flags_cap<<<grid_for(p.dcap), kBlock, 0, st>>>(hd, 0, dels, d.flag.as<uint8_t>());
  kd<<<p.grid_del, kDeltaWaves * kWave, lds_d, st>>>(
      hd, 0, dels, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
      c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), d.roff.as<unsigned long long>(),
      d.deg.as<unsigned long long>(), d.rows.as<uint32_t>(), d.flag.as<uint8_t>(), acc);
  HM_HIP(hipGetLastError());
  // store update into the staging arrays (the removal flags are the deletion mask)
  k_new_sizes_cap<<<grid_for((int64_t)Ecap + 1), kBlock, 0, st>>>(
      hd, Ecap, c.eoff.as<uint32_t>(), d.flag.as<uint8_t>(), ins_off, d.sizes.as<unsigned long long>());
  size_t tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(d.tmp.p, tb, d.sizes.as<unsigned long long>(),
                                          d.noff.as<unsigned long long>(), Ecap + 1, st));
  k_copy_rows_cap<<<grid_for((int64_t)Ecap + 1), kBlock, 0, st>>>(
      hd, Ecap, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), d.flag.as<uint8_t>(), ins_off, ins_nodes,
      d.noff.as<unsigned long long>(), d.nnodes.as<int32_t>(), d.neoff.as<uint32_t>(),
      d.nsize.as<uint8_t>(), d.inc_node.as<int32_t>(), d.inc_edge.as<int32_t>());
  k_pad_inc_cap<<<grid_for(Icap), kBlock, 0, st>>>(hd, Icap, d.inc_node.as<int32_t>(),
                                                    d.inc_edge.as<int32_t>(), (int32_t)((1u << p.bits) - 1u));
  HM_HIP(hipGetLastError());
  tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(d.tmp.p, tb, d.inc_node.as<uint32_t>(), d.srt_node.as<uint32_t>(),
                                            d.inc_edge.as<int32_t>(), d.srt_edge.as<int32_t>(),
                                            (int)Icap, 0, p.bits, st));
  k_offsets_from_sorted<<<grid_for(Icap + 1), kBlock, 0, st>>>(Icap, d.srt_node.as<uint32_t>(), Vcap,
                                                              d.nvoff.as<uint32_t>());
  // the staging arrays become the store (copied: the graph's addresses stay fixed)
  k_copy_words<<<grid_for(Icap), kBlock, 0, st>>>(c.enodes.as<uint32_t>(), d.nnodes.as<uint32_t>(), Icap);
  k_copy_words<<<grid_for((int64_t)Ecap + 1), kBlock, 0, st>>>(c.eoff.as<uint32_t>(), d.neoff.as<uint32_t>(),
                                                                (int64_t)Ecap + 1);
  k_copy_bytes<<<grid_for(Ecap), kBlock, 0, st>>>(c.esize.as<uint8_t>(), d.nsize.as<uint8_t>(), Ecap);
  k_copy_words<<<grid_for(Icap), kBlock, 0, st>>>(c.vedges.as<uint32_t>(), d.srt_edge.as<uint32_t>(), Icap);
  k_copy_words<<<grid_for((int64_t)Vcap + 1), kBlock, 0, st>>>(c.voff.as<uint32_t>(), d.nvoff.as<uint32_t>(),
                                                                (int64_t)Vcap + 1);
  HM_HIP(hipGetLastError());
  // full rows of the graph after the batch (kept for the next update's removal phase)
  k_deg_cap<MT><<<(unsigned)Ecap + 1, kBlock, lds_t, st>>>(hd, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                                           c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                                           d.deg.as<unsigned long long>());
  tb = d.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(d.tmp.p, tb, d.deg.as<unsigned long long>(),
                                          d.roff.as<unsigned long long>(), Ecap + 1, st));
  k_check_rows_cap<<<1, kWave, 0, st>>>(hd, d.roff.as<unsigned long long>(), p.Pcap);
  k_full_fill_cap<MT><<<(unsigned)Ecap, kBlock, lds_t, st>>>(
      hd, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
      d.roff.as<unsigned long long>(), d.rows.as<uint32_t>());
  HM_HIP(hipGetLastError());
  // addition phase: L(inserted ids, graph after the batch)
  k_fill_u8<<<grid_for(Ecap), kBlock, 0, st>>>(d.flag.as<uint8_t>(), Ecap, 0);
  k_set_flags_cap<<<grid_for(p.icap), kBlock, 0, st>>>(hd, 1, dels, d.flag.as<uint8_t>());
  kd<<<p.grid_ins, kDeltaWaves * kWave, lds_d, st>>>(
      hd, 1, dels, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
      c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), d.roff.as<unsigned long long>(),
      d.deg.as<unsigned long long>(), d.rows.as<uint32_t>(), d.flag.as<uint8_t>(), acc);
  HM_HIP(hipGetLastError());
  return HM_OK;
}

// Sizes a plan for a batch (E2 ids, n_del / n_ins changed, Iins inserted incidences, V2 nodes
// after it) with headroom, grows the buffers (keeping the store and the held rows) and captures
// the update. p.cap.exec stays null when the capture cannot be replayed (memset node).
template <typename MT>
static hm_status make_plan(hm_graph* g, int32_t E2, int32_t n_del, int32_t n_ins, int64_t Iins, int32_t V2) {
  DeviceCsr& c = g->csr;
  DynScratch& d = g->dyn;
  UpdPlan& p = g->uplan;
  hipStream_t st = g->stream;
  p.cap.reset();
  p.sig.clear();
  p.wide = sizeof(MT) == 8;
  p.dcap = std::max<int32_t>(64, 2 * n_del);
  p.icap = std::max<int32_t>(64, 2 * n_ins);
  p.iicap = std::max<int64_t>(512, 2 * Iins);
  const int64_t ewant = (int64_t)E2 + std::max<int64_t>({(int64_t)4 * p.icap, (int64_t)E2 / 8, 256});
  p.Ecap = (int32_t)std::min<int64_t>({ewant, (int64_t)dense_max_ids(g, p.wide), kMaxIds - 1});
  if (p.Ecap < E2) return fail(HM_ERR_UNSUPPORTED, "captured update: ids beyond one delta table");
  p.Vcap = (int32_t)std::min<int64_t>((int64_t)V2 + std::max<int64_t>((int64_t)V2 / 8, p.iicap), (1ll << 30));
  if (p.Vcap < V2) return fail(HM_ERR_UNSUPPORTED, "captured update: node ids too large");
  p.Icap = c.I + Iins + std::max<int64_t>(c.I / 8, 4 * p.iicap);
  if (p.Icap >= (1ll << 31)) return fail(HM_ERR_UNSUPPORTED, "captured update: too many incidences");
  p.Pcap = d.rows_P + d.rows_P / 4 + 4096;
  p.bits = 1;
  while (p.bits < 31 && ((int64_t)1 << p.bits) <= (int64_t)p.Vcap) ++p.bits;
  p.grid_del = delta_grid_cap(p.dcap);
  p.grid_ins = delta_grid_cap(p.icap);
  const size_t E1 = (size_t)p.Ecap + 1, V1 = (size_t)p.Vcap + 1, Ic = (size_t)p.Icap;
  // the store and the held rows keep their contents
  HM_TRY(grow_keep(c.eoff, 4 * E1, st));
  HM_TRY(grow_keep(c.enodes, 4 * Ic, st));
  HM_TRY(grow_keep(c.esize, E1, st));
  HM_TRY(grow_keep(c.voff, 4 * V1, st));
  HM_TRY(grow_keep(c.vedges, 4 * Ic, st));
  HM_TRY(grow_keep(d.deg, 8 * E1, st));
  HM_TRY(grow_keep(d.roff, 8 * E1, st));
  HM_TRY(grow_keep(d.rows, 4 * (size_t)p.Pcap, st));
  HM_TRY(d.flag.ensure(E1));
  HM_TRY(d.acc.ensure(sizeof(unsigned long long) * kDynAccCount));
  HM_TRY(d.sizes.ensure(8 * E1));
  HM_TRY(d.noff.ensure(8 * E1));
  HM_TRY(d.neoff.ensure(4 * E1));
  HM_TRY(d.nsize.ensure(E1));
  for (DevBuf* b : {&d.nnodes, &d.inc_node, &d.inc_edge, &d.srt_node, &d.srt_edge}) HM_TRY(b->ensure(4 * Ic));
  HM_TRY(d.nvoff.ensure(4 * V1));
  size_t t1 = 0, t3 = 0;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, d.sizes.as<unsigned long long>(),
                                          d.noff.as<unsigned long long>(), (int)E1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t3, d.inc_node.as<uint32_t>(), d.srt_node.as<uint32_t>(),
                                            d.inc_edge.as<int32_t>(), d.srt_edge.as<int32_t>(),
                                            (int)Ic, 0, p.bits, st));
  HM_TRY(d.tmp.ensure(std::max({t1, t3, (size_t)16})));
  const InLayout L = in_layout(p);
  HM_TRY(p.in.ensure(L.total));
  HM_TRY(p.hin.ensure(L.total));
  HM_TRY(p.hout.ensure(kUpdOutBytes));
  // nothing may allocate or set attributes while the stream is captured
  HM_TRY(allow_lds((const void*)k_delta_cap<MT, kDeltaWaves, kDeltaLoadsWide>, delta_lds<MT>(p.Ecap)));
  HM_TRY(allow_lds((const void*)k_delta_cap<MT, kDeltaWaves, kDeltaLoadsNarrow>, delta_lds<MT>(p.Ecap)));
  HM_TRY(allow_lds((const void*)k_deg_cap<MT>, table_lds<MT>(p.Ecap)));
  HM_TRY(allow_lds((const void*)k_full_fill_cap<MT>, table_lds<MT>(p.Ecap)));
  CaptureGuard capturing;  // no allocation or free on another thread meanwhile
  HM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const hm_status s1 = enqueue_update_cap<MT>(g);
  CapturedGraph cap;
  const hm_status s2 = end_capture(st, &cap);  // ends the capture on every path
  if (s1 != HM_OK || s2 != HM_OK) {
    cap.reset();
    return s1 != HM_OK ? s1 : s2;
  }
  p.cap = cap;
  p.sig = plan_sig(g);
  ++p.captures;
  return HM_OK;
}

// The update as one captured graph. *handled = false (and nothing changed but buffer capacities
// and the held rows) when the plan cannot be made: the caller runs the eager update.
template <typename MT>
static hm_status update_captured(hm_graph* g, const std::vector<int32_t>& dels_sorted,
                                 const std::vector<int64_t>& ioff, const std::vector<int32_t>& inodes,
                                 int32_t n_ins, int32_t newV, unsigned long long* h, bool* handled) {
  *handled = false;
  DeviceCsr& c = g->csr;
  DynScratch& d = g->dyn;
  UpdPlan& p = g->uplan;
  hipStream_t st = g->stream;
  const int32_t E = c.E, E2 = E + n_ins, n_del = (int32_t)dels_sorted.size();
  const int64_t Iins = ioff.back();
  if (!d.rows_valid || d.rows_tile != 0) {  // the removal phase reads the rows of the graph as it is
    d.rows_valid = false;
    HM_TRY(build_full_rows<MT>(g, d, 0));
    d.rows_valid = true;
    d.rows_tile = 0;
  }
  const bool fits = p.cap.exec && p.wide == (sizeof(MT) == 8) && E2 <= p.Ecap && n_del <= p.dcap &&
                    n_ins <= p.icap && Iins <= p.iicap && newV <= p.Vcap && c.I + Iins <= p.Icap &&
                    p.sig == plan_sig(g);
  if (!fits) {
    if (make_plan<MT>(g, E2, n_del, n_ins, Iins, newV) != HM_OK) {
      clear_error();
      p.cap.reset();
      p.disabled = true;  // this graph's updates stay eager
      return HM_OK;
    }
    if (!p.cap.exec) {
      p.disabled = true;  // the capture held a memset node
      return HM_OK;
    }
  }
  const InLayout L = in_layout(p);
  // the staging block, the device copy of it and the write-back must lie inside the buffers the
  // plan sized (a batch that fits the plan fits these; anything else is an engine bug)
  if (n_del > p.dcap || n_ins > p.icap || Iins > p.iicap || (int64_t)ioff.size() != (int64_t)n_ins + 1 ||
      (int64_t)inodes.size() < Iins || L.total > p.hin.bytes || L.total > p.in.bytes ||
      kUpdOutBytes > p.hout.bytes || !p.hin.p || !p.hout.p || !p.in.p ||
      L.ins_off + sizeof(uint32_t) * ((size_t)n_ins + 1) > L.ins_nodes ||
      L.ins_nodes + sizeof(int32_t) * (size_t)Iins > L.total ||
      L.dels + sizeof(int32_t) * (size_t)n_del > L.ins_off)
    return fail(HM_ERR_DEVICE, "captured update: staging layout does not hold the batch");
  unsigned char* hin = p.hin.as<unsigned char>();
  UpdHdr hdr{};
  hdr.E_old = E;
  hdr.E_new = E2;
  hdr.n_del = n_del;
  hdr.n_ins = n_ins;
  hdr.V_new = newV;
  hdr.parts_del = delta_parts(n_del);
  hdr.parts_ins = delta_parts(n_ins);
  std::memcpy(hin, &hdr, sizeof(hdr));
  if (n_del) std::memcpy(hin + L.dels, dels_sorted.data(), sizeof(int32_t) * (size_t)n_del);
  uint32_t* io = reinterpret_cast<uint32_t*>(hin + L.ins_off);
  for (int32_t k = 0; k <= n_ins; ++k) io[k] = (uint32_t)ioff[(size_t)k];
  if (Iins) std::memcpy(hin + L.ins_nodes, inodes.data(), sizeof(int32_t) * (size_t)Iins);
  *handled = true;
  d.rows_valid = false;  // rebuilt by the launch unless they overflow
  struct DrainOnExit {
    hipStream_t s;
    ~DrainOnExit() { (void)hipStreamSynchronize(s); }
  } drain{st};
  HM_HIP(hipMemcpyAsync(p.in.p, hin, L.total, hipMemcpyHostToDevice, st));
  HM_HIP(hipEventRecord(g->ev[3], st));
  HM_HIP(hipGraphLaunch(p.cap.exec, st));
  HM_HIP(hipEventRecord(g->ev[7], st));
  unsigned char* out = p.hout.as<unsigned char>();
  HM_HIP(hipMemcpyAsync(out, d.acc.p, sizeof(unsigned long long) * kDynAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(out + sizeof(unsigned long long) * kDynAccCount, p.in.p, sizeof(UpdHdr),
                        hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  UpdHdr back;
  std::memcpy(&back, out + sizeof(unsigned long long) * kDynAccCount, sizeof(back));
  c.E = E2;
  c.V = newV;
  c.I = (int64_t)back.I_new;
  float ms = 0.f;
  HM_HIP(hipEventElapsedTime(&ms, g->ev[3], g->ev[7]));
  g->ms_upd_rows = 0.f;
  g->ms_upd_store = 0.f;
  g->ms_upd_delta = ms;
  p.last_captured = true;
  p.last_overflow = back.overflow != 0;
  if (!back.overflow) {
    d.rows_valid = true;
    d.rows_tile = 0;
    d.rows_P = back.P_new;
    std::memcpy(h, out, sizeof(unsigned long long) * kDynAccCount);
    return HM_OK;
  }
  // rows over the plan: the device acc holds the removal phase; the addition phase runs eagerly
  // (larger row buffers; the next update captures a larger plan)
  std::vector<int32_t> ins((size_t)n_ins);
  for (int32_t k = 0; k < n_ins; ++k) ins[(size_t)k] = E + k;
  HM_TRY(local_count_any(g, d, ins.data(), n_ins, +1, 4));
  HM_HIP(hipMemcpyAsync(h, d.acc.p, sizeof(unsigned long long) * kDynAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  if (n_ins > 0) HM_TRY(add_phase_ms(g, 4));
  return HM_OK;
}

}  // namespace hm

using namespace hm;

extern "C" {

// Update with an exact delta by the hub-aware parts (graphs whose frequent nodes put the full
// projection rows k_delta needs out of reach): (A) + (B) + (C) of the graph before and after the
// batch, and (D) over the light triangles that hold a deleted hyperedge (before) or an inserted
// one (after) — the changed ids are ranked first in the light order, so (D) over those centres
// alone enumerates exactly them, and the light triangles without a changed member are common to
// both graphs and cancel. Both graphs use one hub set: the one chosen on the graph before.
static hm_status update_hub(hm_graph* g, DynScratch& d, const int32_t* del_ids, int32_t n_del,
                            const std::vector<uint8_t>& delmask, const std::vector<uint32_t>& ins_off32,
                            const std::vector<int32_t>& inodes, int32_t n_ins, int32_t newV, int32_t maxsz,
                            int32_t* assigned_ids, int64_t delta_out[HM_STATIC_BINS]) {
  const int32_t E = g->csr.E;
  std::vector<int32_t> dels(del_ids, del_ids + n_del);
  std::sort(dels.begin(), dels.end());
  int32_t h = g->hubs;
  if (h < 0) {  // the automatic hub count is chosen once and kept across updates (a batch changes
                // the graph by a few percent; the cost model reads the whole store back)
    if (g->hub_auto_h <= 0) {
      HM_TRY(hub_choose(g, &h, nullptr));
      g->hub_auto_h = h;
    }
    // A delta redoes the class, pair-type and light-wedge parts of both graphs, which grow with
    // the hub count, and only the light triangles of the changed hyperedges, which shrink with it:
    // fewer hubs than a count's choice (one batch of the 500k-hyperedge dynamic workload: 8.2 s with
    // 8 hubs, 5.7 s with 12, 15.4 s with 16, 25.0 s with the 23 a count picks).
    h = std::min<int32_t>(g->hub_auto_h, kHubDeltaHubs);
  }
  std::vector<int32_t> hubs;
  unsigned long long abc0[kAccCount], d0[kAccCount], abc1[kAccCount], d1[kAccCount];
  // the hub parts read their results back before returning: host time is their time
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  for (;;) {  // graph before the batch; an overflowing hub count is lowered (automatic mode)
    HubRun r0;
    r0.hubs_out = &hubs;
    r0.first = dels.data();
    r0.n_first = n_del;
    r0.only_first = true;
    r0.stats = false;
    bool overflow = false;
    const hm_status st = hub_parts(g, h, r0, abc0, d0, &overflow);
    if (st == HM_OK) break;
    if (g->hubs > 0 || !overflow || h == 1) return st;
    h = std::max<int32_t>(1, h * 3 / 4);
    g->hub_auto_h = std::min(g->hub_auto_h, h);
    clear_error();
  }
  const double ms_before = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  ++g->version;
  g->hub_auto_version = g->version;  // the kept hub count stays valid for the graph after the batch
  d.rows_valid = false;  // the held rows (k_delta's) describe no graph after this update
  HM_HIP(hipEventRecord(g->ev[6], g->stream));  // the hub parts use events 0..5
  HM_TRY(apply_batch(g, d, delmask, ins_off32, inodes, n_ins, newV));
  HM_HIP(hipEventRecord(g->ev[7], g->stream));
  g->csr.max_size = maxsz;
  for (int32_t k = 0; k < n_del; ++k) g->live[(size_t)del_ids[k]] = 0;
  g->live.resize((size_t)E + n_ins, 1);
  for (int32_t k = 0; k < n_ins; ++k) assigned_ids[k] = E + k;
  std::vector<int32_t> ins((size_t)n_ins);
  for (int32_t k = 0; k < n_ins; ++k) ins[(size_t)k] = E + k;
  const auto t1 = clk::now();
  HubRun r1;
  r1.hub_nodes = &hubs;
  r1.first = ins.data();
  r1.n_first = n_ins;
  r1.only_first = true;
  r1.stats = false;
  bool overflow = false;
  const hm_status st = hub_parts(g, (int32_t)hubs.size(), r1, abc1, d1, &overflow);
  if (st != HM_OK) {
    set_error("hm_graph_update: the batch is applied, but its hub-aware delta failed on the graph after it (%s); "
              "recount the graph", overflow ? "a neighbour-key histogram overflowed" : "see the earlier error");
    return st;
  }
  for (int k = 0; k < HM_STATIC_BINS; ++k)
    delta_out[k] = (int64_t)((abc1[kAccHist + k] - abc0[kAccHist + k]) + (d1[kAccHist + k] - d0[kAccHist + k]));
  g->upd_removed = (int64_t)d0[kAccLightTri];  // light triangles enumerated before / after the batch
  g->upd_added = (int64_t)d1[kAccLightTri];
  g->upd_entries = -1;
  g->ms_upd_delta = (float)(ms_before + std::chrono::duration<double, std::milli>(clk::now() - t1).count());
  HM_HIP(hipEventSynchronize(g->ev[7]));
  HM_HIP(hipEventElapsedTime(&g->ms_upd_store, g->ev[6], g->ev[7]));
  g->have_stats = g->stats_lazy = false;
  return HM_OK;
}

hm_status hm_graph_update(hm_graph* g, const int32_t* del_ids, int32_t n_del,
                          const int64_t* ins_offsets, const int32_t* ins_nodes, int32_t n_ins,
                          int32_t* assigned_ids, int64_t delta_out[HM_STATIC_BINS]) {
  clear_error();
  if (!g) return fail(HM_ERR_BAD_INPUT, "hm_graph_update: null graph");
  if (n_del < 0 || n_ins < 0) return fail(HM_ERR_BAD_INPUT, "hm_graph_update: negative count");
  if ((n_del > 0 && !del_ids) || (n_ins > 0 && (!ins_offsets || !assigned_ids)))
    return fail(HM_ERR_BAD_INPUT, "hm_graph_update: null batch array");
  const int32_t E = g->csr.E;
  if ((int64_t)E + n_ins >= kMaxIds) return fail(HM_ERR_OVER_BUDGET, "hm_graph_update: id space exhausted");
  std::vector<uint8_t> delmask((size_t)std::max(E, 1), 0);
  for (int32_t k = 0; k < n_del; ++k) {
    const int32_t x = del_ids[k];
    if (x < 0 || x >= E) return fail(HM_ERR_BAD_INPUT, "hm_graph_update: deleted id out of range");
    if (!g->live[(size_t)x]) return fail(HM_ERR_BAD_INPUT, "hm_graph_update: deleted id is not live");
    if (delmask[(size_t)x]) return fail(HM_ERR_BAD_INPUT, "hm_graph_update: deleted id repeated");
    delmask[(size_t)x] = 1;
  }
  std::vector<int64_t> ioff;
  std::vector<int32_t> inodes;
  int32_t vmax = -1;
  hm_status st = prepare_host_rows(ins_offsets, ins_nodes, n_ins, &ioff, &inodes, &vmax);
  if (st != HM_OK) return st;
  std::vector<uint32_t> ins_off32(ioff.size());
  int32_t maxsz = g->csr.max_size;
  for (size_t k = 0; k < ioff.size(); ++k) ins_off32[k] = (uint32_t)ioff[k];
  for (int32_t k = 0; k < n_ins; ++k) maxsz = std::max<int32_t>(maxsz, (int32_t)(ioff[(size_t)k + 1] - ioff[(size_t)k]));
  const int32_t newV = std::max(g->csr.V, vmax + 1);
  const bool want = delta_out != nullptr;
  // a delta needs masks of every changed hyperedge: refuse before the store changes
  if (want && maxsz > HM_MAX_EDGE_SIZE)
    return fail(HM_ERR_UNSUPPORTED, "hm_graph_update: a delta needs hyperedges of at most 64 nodes");
  HM_HIP(hipSetDevice(g->device));
  DynScratch& d = dyn_scratch(g);
  if (want && g->hubs != 0) {
    g->ms_upd_rows = g->ms_upd_delta = 0.f;
    g->upd_removed = g->upd_added = g->upd_entries = -1;
    g->uplan.last_captured = g->uplan.last_overflow = false;
    return update_hub(g, d, del_ids, n_del, delmask, ins_off32, inodes, n_ins, newV, maxsz, assigned_ids,
                      delta_out);
  }
  ++g->version;  // a captured count of the old graph is never replayed
  g->ms_upd_rows = g->ms_upd_delta = 0.f;
  g->upd_removed = g->upd_added = g->upd_entries = -1;
  g->uplan.last_captured = g->uplan.last_overflow = false;
  if (want && g->uplan.on && !g->uplan.disabled && g->tile == 0 && dense_fits(g, E + n_ins, maxsz > 32)) {
    std::vector<int32_t> ds(del_ids, del_ids + n_del);
    std::sort(ds.begin(), ds.end());
    const int32_t old_max = g->csr.max_size;
    g->csr.max_size = maxsz;  // one mask width for both phases
    unsigned long long h[kDynAccCount];
    bool handled = false;
    const hm_status r = maxsz > 32
                            ? update_captured<mask64_t>(g, ds, ioff, inodes, n_ins, newV, h, &handled)
                            : update_captured<uint32_t>(g, ds, ioff, inodes, n_ins, newV, h, &handled);
    if (!handled) g->csr.max_size = old_max;
    if (r != HM_OK) return r;
    if (handled) {
      for (int32_t k = 0; k < n_del; ++k) g->live[(size_t)del_ids[k]] = 0;
      g->live.resize((size_t)E + n_ins, 1);
      for (int32_t k = 0; k < n_ins; ++k) assigned_ids[k] = E + k;
      for (int k = 0; k < HM_STATIC_BINS; ++k) delta_out[k] = (int64_t)h[k];
      g->upd_removed = (int64_t)h[kDynAccRemoved];
      g->upd_added = (int64_t)h[kDynAccAdded];
      g->upd_entries = (int64_t)h[kDynAccEntries];
      g->have_stats = g->stats_lazy = false;
      return HM_OK;
    }
  }
  // host ids copied to the device by the removal phase: kept alive until the stream has drained
  // (the guard, destroyed first, waits on every return path)
  std::vector<int32_t> dels(del_ids, del_ids + n_del);
  struct DrainOnExit {
    hipStream_t s;
    ~DrainOnExit() { (void)hipStreamSynchronize(s); }
  } drain{g->stream};
  if (want) {
    HM_TRY(d.acc.ensure(sizeof(unsigned long long) * kDynAccCount));
    k_zero_acc<<<1, kDynAccCount, 0, g->stream>>>(d.acc.as<unsigned long long>(), kDynAccCount);
    std::sort(dels.begin(), dels.end());
    HM_TRY(local_count_any(g, d, dels.data(), n_del, -1, 0));
  }
  d.rows_valid = false;  // the store changes: the held rows describe the old graph
  HM_HIP(hipEventRecord(g->ev[3], g->stream));
  HM_TRY(apply_batch(g, d, delmask, ins_off32, inodes, n_ins, newV));
  HM_HIP(hipEventRecord(g->ev[7], g->stream));  // read once the stream has drained
  g->csr.max_size = maxsz;
  for (int32_t k = 0; k < n_del; ++k) g->live[(size_t)del_ids[k]] = 0;
  g->live.resize((size_t)E + n_ins, 1);
  for (int32_t k = 0; k < n_ins; ++k) assigned_ids[k] = E + k;
  if (want) {
    std::vector<int32_t> ins((size_t)n_ins);
    for (int32_t k = 0; k < n_ins; ++k) ins[(size_t)k] = E + k;
    HM_TRY(local_count_any(g, d, ins.data(), n_ins, +1, 4));
    unsigned long long h[kDynAccCount];
    HM_HIP(hipMemcpyAsync(h, d.acc.p, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    HM_HIP(hipStreamSynchronize(g->stream));
    if (n_del > 0) HM_TRY(add_phase_ms(g, 0));
    if (n_ins > 0) HM_TRY(add_phase_ms(g, 4));
    for (int k = 0; k < HM_STATIC_BINS; ++k) delta_out[k] = (int64_t)h[k];
    g->upd_removed = (int64_t)h[kDynAccRemoved];
    g->upd_added = (int64_t)h[kDynAccAdded];
    g->upd_entries = (int64_t)h[kDynAccEntries];
  }
  HM_HIP(hipEventSynchronize(g->ev[7]));
  HM_HIP(hipEventElapsedTime(&g->ms_upd_store, g->ev[3], g->ev[7]));
  g->have_stats = g->stats_lazy = false;
  return HM_OK;
}

hm_status hm_graph_update_stats(hm_graph* g, hm_update_stats* out) {
  clear_error();
  if (!g || !out) return fail(HM_ERR_BAD_INPUT, "hm_graph_update_stats: null argument");
  if (g->upd_removed < 0) return fail(HM_ERR_BAD_INPUT, "hm_graph_update_stats: no update with a delta has run");
  out->removed = g->upd_removed;
  out->added = g->upd_added;
  out->row_entries = g->upd_entries;
  out->ms_rows = g->ms_upd_rows;
  out->ms_delta = g->ms_upd_delta;
  out->ms_store = g->ms_upd_store;
  out->captured = g->uplan.last_captured ? 1 : 0;
  out->rows_overflow = g->uplan.last_overflow ? 1 : 0;
  return HM_OK;
}

hm_status hm_graph_set_update_capture(hm_graph* g, int32_t on) {
  clear_error();
  if (!g) return fail(HM_ERR_BAD_INPUT, "hm_graph_set_update_capture: null handle");
  UpdPlan& p = g->uplan;
  p.on = on != 0;
  if (!p.on) {
    p.cap.reset();
    p.sig.clear();
    p.disabled = false;
  }
  return HM_OK;
}

hm_status hm_graph_update_capture_state(hm_graph* g, hm_update_capture_state* out) {
  clear_error();
  if (!g || !out) return fail(HM_ERR_BAD_INPUT, "hm_graph_update_capture_state: null argument");
  const UpdPlan& p = g->uplan;
  *out = hm_update_capture_state{};
  out->on = p.on ? 1 : 0;
  out->disabled = p.disabled ? 1 : 0;
  out->captured = p.cap.exec ? 1 : 0;
  out->captures = p.captures;
  out->last_captured = p.last_captured ? 1 : 0;
  out->last_overflow = p.last_overflow ? 1 : 0;
  out->id_capacity = p.Ecap;
  out->node_capacity = p.Vcap;
  out->incidence_capacity = p.Icap;
  out->row_capacity = (int64_t)p.Pcap;
  out->batch_deletes = p.dcap;
  out->batch_inserts = p.icap;
  out->batch_incidences = p.iicap;
  return HM_OK;
}

hm_status hm_count_delta(hm_graph* old_g, hm_graph* new_g, const int32_t* del_ids, int32_t n_del,
                         const int32_t* ins_ids, int32_t n_ins, int64_t delta_out[HM_STATIC_BINS]) {
  clear_error();
  if (!old_g || !new_g || !delta_out) return fail(HM_ERR_BAD_INPUT, "hm_count_delta: null argument");
  if (old_g == new_g) return fail(HM_ERR_BAD_INPUT, "hm_count_delta: old and new must be two snapshots");
  if (n_del < 0 || n_ins < 0 || (n_del > 0 && !del_ids) || (n_ins > 0 && !ins_ids))
    return fail(HM_ERR_BAD_INPUT, "hm_count_delta: bad id list");
  std::vector<int32_t> dels(del_ids, del_ids + n_del), ins(ins_ids, ins_ids + n_ins);
  std::sort(dels.begin(), dels.end());
  std::sort(ins.begin(), ins.end());
  if (std::adjacent_find(dels.begin(), dels.end()) != dels.end() ||
      std::adjacent_find(ins.begin(), ins.end()) != ins.end())
    return fail(HM_ERR_BAD_INPUT, "hm_count_delta: repeated id");
  for (int32_t x : dels)
    if (x < 0 || x >= old_g->csr.E) return fail(HM_ERR_BAD_INPUT, "hm_count_delta: deleted id out of range");
  for (int32_t x : ins)
    if (x < 0 || x >= new_g->csr.E) return fail(HM_ERR_BAD_INPUT, "hm_count_delta: inserted id out of range");
  unsigned long long h[32] = {0};
  for (hm_graph* g : {old_g, new_g}) {
    HM_HIP(hipSetDevice(g->device));
    DynScratch& d = dyn_scratch(g);
    HM_TRY(d.acc.ensure(sizeof(unsigned long long) * kDynAccCount));
    k_zero_acc<<<1, kDynAccCount, 0, g->stream>>>(d.acc.as<unsigned long long>(), kDynAccCount);
    const bool is_old = g == old_g;
    HM_TRY(local_count_any(g, d, is_old ? dels.data() : ins.data(), is_old ? n_del : n_ins,
                           is_old ? -1 : +1, -1));
    unsigned long long part[32];
    HM_HIP(hipMemcpyAsync(part, d.acc.p, sizeof(part), hipMemcpyDeviceToHost, g->stream));
    HM_HIP(hipStreamSynchronize(g->stream));
    for (int k = 0; k < 32; ++k) h[k] += part[k];
  }
  for (int k = 0; k < HM_STATIC_BINS; ++k) delta_out[k] = (int64_t)h[k];
  return HM_OK;
}

}  // extern "C"
