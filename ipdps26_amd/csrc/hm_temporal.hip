// Temporal h-motif counting (THyMe's 96 temporal h-motifs) on MI355X (gfx950).
//
// Semantics follow THyMe+ (THyMe-main/code/main_thymeP.cpp:54-389): a temporal h-motif instance is
// three temporal hyperedges whose time span is at most delta (window kept inclusive,
// main_thymeP.cpp:228) and of which at least two pairs overlap; it is classified by the Venn
// pattern of the three node sets written in time order (THyMe-main/code/motif_id.cpp:57-108).
// Temporal hyperedges with identical node sets are copies of one static hyperedge; THyMe+ counts
// their combinations through per-static-edge time lists (instances :299-334, motifs 87-95
// :338-363, motif 96 :366-367). Counting triples of temporal hyperedges directly, as done here,
// gives the same numbers (a copy pair is just a pair whose overlap is the whole node set).
//
// Layout: temporal hyperedges in time order (ties by input order) are "positions".
//   H2V by position: eoff/enodes/esize; V2H: voff/vpos (positions ascending); times[E].
//   Window of anchor i: positions [lo_i, i) with t >= t_i - delta.
//   Temporal projection rows (built per count): F(x) = positions y > x with t_y <= t_x + delta
//   that overlap x, B(x) = positions y < x with t_y >= t_x - delta that overlap x; entries packed
//   as (y << 7 | |x ∩ y|), ascending.
//
// Counting kernel (one workgroup per anchor i; every instance is counted at its newest edge i):
//   M[y] = bitmask of i's nodes contained in y, for y in the window (LDS, dense over the window),
//   with W16[y] = (|y ∩ i|, |y \ i|); the neighbours of i are its backward row B(i).
//   open at i:  every ordered pair (x, y) of B(i) counted as if x ∩ y = ∅, from class histograms
//               (ordered_open_pairs: C with i, [node outside i], [C_x + C_y < |i|]);
//   for every x in B(i) (one wave per x):
//     (b) y in F(x), y < i:   y a neighbour of i: the closed triangle (x, y, i) +1 in its bin,
//                             -1 in the as-if-open bin counted above and -1 in its middle edge's
//                             side (e) bin (below); else (x, y, i) is open at x.
// Side (e), triples (y, x, i) open at x with y earlier (k_twedge, one wave per x): every pair
// (y in B(x), i in F(x)) with y >= lo_i counted by class whether or not y overlaps i; the pairs
// where it does are the triangles, taken back by the counting kernel as above.
// Windows longer than one LDS table use the tiled counting kernel below (position tiles, the
// anchor's neighbours read from its own backward row) with the same k_twedge.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <numeric>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_tables.hpp"

namespace hm {

__constant__ int8_t c_temporal_bin[128] = HM_TEMPORAL_BIN_TABLE;

constexpr int kTemporalBins = 96;
enum TAccSlot : int {
  kTAccHist = 0,        // 96 bins
  kTAccPairs = 96,      // sum_x |F(x)| (temporal projection pairs)
  kTAccMaxWin = 97,     // max_i (i - lo_i)
  kTAccNbrPairs = 98,   // sum_i C(|N(i) ∩ window|, 2)
  kTAccOrient = 99,     // sum_x sum_{y in F(x)} d_y  (d = fdeg + bdeg, the delta-window degree)
  kTAccCount = 104
};

// lower_bound over an ascending array
template <typename T>
__device__ inline uint32_t lower_bound_dev(const T* a, uint32_t n, T key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Window bounds: lo[x] = first position with t >= t_x - delta, hi[x] = last with t <= t_x + delta.
__global__ void k_twindow(int32_t E, const long long* __restrict__ times, long long delta,
                          int32_t* __restrict__ lo, int32_t* __restrict__ hi) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= E) return;
  const long long t = times[x];
  lo[x] = (int32_t)lower_bound_dev<long long>(times, (uint32_t)x + 1, t - delta);
  // first position with time > t + delta, minus one
  uint32_t l = (uint32_t)x, h = (uint32_t)E;
  while (l < h) {
    uint32_t mid = (l + h) >> 1;
    if (times[mid] <= t + delta) l = mid + 1;
    else h = mid;
  }
  hi[x] = (int32_t)l - 1;
}

// M[y - base] |= bit(k) for node k of hyperedge x and every position y in [base, last] holding it,
// y != x. Positions per node are ascending, so the range is found by binary search.
template <typename MT>
__device__ inline void build_window_masks(MT* M, int32_t x, int32_t base, int32_t last,
                                          const uint32_t* __restrict__ eoff,
                                          const int32_t* __restrict__ enodes,
                                          const uint32_t* __restrict__ voff,
                                          const int32_t* __restrict__ vpos) {
  const int32_t R = last - base + 1;
  for (int k = threadIdx.x; k < R; k += blockDim.x) M[k] = (MT)0;
  __syncthreads();
  const uint32_t s = eoff[x], t = eoff[x + 1];
  for (uint32_t i = s; i < t; ++i) {
    const int32_t v = enodes[i];
    const MT bit = (MT)1 << (i - s);
    const int32_t* row = vpos + voff[v];
    const uint32_t n = voff[v + 1] - voff[v];
    const uint32_t k0 = lower_bound_dev<int32_t>(row, n, base);
    const uint32_t k1 = lower_bound_dev<int32_t>(row, n, last + 1);
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
      const int32_t y = row[k];
      if (y != x) atomicOr(&M[y - base], bit);
    }
  }
  __syncthreads();
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_tproj_count(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vpos,
    const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    unsigned long long* __restrict__ fdeg, unsigned long long* __restrict__ bdeg) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t red[2][kWavesPerBlock];
  const int32_t x = blockIdx.x;
  const int32_t base = lo[x], last = hi[x];
  build_window_masks<MT>(M, x, base, last, eoff, enodes, voff, vpos);
  uint32_t f = 0, b = 0;
  for (int32_t y = base + (int32_t)threadIdx.x; y <= last; y += blockDim.x) {
    if (M[y - base]) {
      if (y > x) ++f;
      else ++b;
    }
  }
  f = wave_sum(f);
  b = wave_sum(b);
  if (lane_id() == 0) {
    red[0][wave_id()] = f;
    red[1][wave_id()] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ff = 0, bb = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
      ff += red[0][w];
      bb += red[1][w];
    }
    fdeg[x] = ff;
    bdeg[x] = bb;
  }
}

// Compacts positions y in [from, to] with M[y - base] != 0 into out[] (ascending).
template <typename MT>
__device__ inline uint32_t compact_rows(const MT* M, int32_t base, int32_t from, int32_t to,
                                        uint32_t* __restrict__ out, uint32_t* wcnt) {
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t running = 0;
  for (int32_t y0 = from; y0 <= to; y0 += blockDim.x) {
    const int32_t y = y0 + (int32_t)threadIdx.x;
    const MT m = (y <= to) ? M[y - base] : (MT)0;
    const unsigned long long bal = __ballot(m != (MT)0);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) {
      woff += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    if (m) out[running + woff + (uint32_t)__popcll(bal & below)] =
        ((uint32_t)y << kCntBits) | (uint32_t)popc(m);
    running += tot;
    __syncthreads();
  }
  return running;
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_tproj_fill(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vpos,
    const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    const unsigned long long* __restrict__ foff, const unsigned long long* __restrict__ boff,
    uint32_t* __restrict__ frows, uint32_t* __restrict__ brows) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t x = blockIdx.x;
  const int32_t base = lo[x], last = hi[x];
  build_window_masks<MT>(M, x, base, last, eoff, enodes, voff, vpos);
  compact_rows<MT>(M, base, x + 1, last, frows + foff[x], wcnt);
  compact_rows<MT>(M, base, base, x - 1, brows + boff[x], wcnt);
}

// Layout of the counting kernel's dynamic LDS for a window of W positions: the anchor's masks
// M[W] and, per position, W16 = C_yi | (s_y - C_yi) << 8 (overlap with the anchor, nodes outside
// it).
template <typename MT>
struct TCountLds {
  static size_t bytes(int32_t W) {
    size_t b = ((size_t)W * sizeof(MT) + 15) & ~(size_t)15;  // M
    b += ((size_t)W * sizeof(uint16_t) + 15) & ~(size_t)15;  // W16
    return b;
  }
};

// As-if-open bin of the ordered pair (x earlier, y later) of neighbours of anchor i, i.e. the bin
// the triple (x, y, i) would have if x ∩ y = ∅: Venn id 48 + [x ⊄ i] + 2 [y ⊄ i] + 4 [i ⊄ x ∪ y].
__device__ inline int as_open_id(int fx, int fy, int cxi, int cyi, int si) {
  return 48 + fx + 2 * fy + 4 * (int)(cxi + cyi < si);
}

// Slots of closed triangles (x, y, i) in time order. A triangle adds +1 to the bin of its Venn id
// v (bits: 0-2 the private regions of x, y, i; 3-5 the pairwise-only regions xy, yi, xi; 6 the
// triple region), -1 to its as-if-open bin (the pair x, y was counted as open at i) and -1 to
// its middle edge's side (e) bin (the pair (x, i) of y, counted by k_twedge). The as-if-open bin
// is fixed by v and b = [C_xi + C_yi < s_i] (x ⊄ i is Venn bit 0 or 3, y ⊄ i bit 1 or 3), the
// side (e) bin Y' | X' << 1 | I' << 2 | 8 | 16 by v and b2 = X' = [|y| > C_xy + C_yi]
// (Y' = [x ⊄ y] is bit 0 or 5, I' = [i ⊄ y] bit 2 or 5). Without a triple region b is bit 2 and
// b2 bit 1, and every pairwise-only region is non-empty. So one counter per (v, b, b2) holds all
// three: slot = b2 | b << 1 | low << 2 | mid << 5 with x ∩ y ∩ i non-empty (0..255), 264 + low
// without (264..271), low = bits 0-2 and mid = bits 3-5 of v. Slots 256..263 count the triples
// (x, y, i) open at x with y between x and i (y ∩ i = ∅) by X | Y << 1 | I << 2, X = [x has a
// node in neither y nor i], Y = [y ⊄ x], I = [i ⊄ x]: Venn id X | Y << 1 | I << 2 | 8 | 32.
constexpr int kTSlots = 272;
constexpr int kTSlotOpenX = 256;
constexpr int kTSlotNoTriple = 264;

// Venn id, as-if-open id and side (e) id of the middle edge, of triangle slot p (p < 256 or
// p >= kTSlotNoTriple).
__device__ inline void tslot_bins(int p, int* v, int* as_open, int* side_e) {
  int vv, b, b2;
  if (p < 256) {
    const int q = p >> 1;
    b2 = p & 1;
    b = q & 1;
    vv = ((q >> 1) & 7) | (q >> 4) << 3 | 64;
  } else {
    vv = (p & 7) | 56;
    b = (vv >> 2) & 1;
    b2 = (vv >> 1) & 1;
  }
  const int fx = (vv & 1) | ((vv >> 3) & 1), fy = ((vv >> 1) & 1) | ((vv >> 3) & 1);
  const int ye = (vv & 1) | ((vv >> 5) & 1), ie = ((vv >> 2) & 1) | ((vv >> 5) & 1);
  *v = vv;
  *as_open = 48 + fx + 2 * fy + 4 * b;  // as_open_id with [C_xi + C_yi < s_i] = b
  *side_e = ye | b2 << 1 | ie << 2 | 8 | 16;
}

// Adds the count s of slot p into the bin sums.
__device__ inline void tslot_fold(int p, unsigned long long s, unsigned long long* bsum,
                                  const int8_t* tbl) {
  if (p >= kTSlotOpenX && p < kTSlotNoTriple) {
    atomicAdd(&bsum[tbl[(p - kTSlotOpenX) | 8 | 32]], s);
  } else {
    int v, as_open, side_e;
    tslot_bins(p, &v, &as_open, &side_e);
    const unsigned long long neg = (unsigned long long)(-(long long)s);
    atomicAdd(&bsum[tbl[v]], s);
    atomicAdd(&bsum[tbl[as_open]], neg);
    atomicAdd(&bsum[tbl[side_e]], neg);
  }
}

// Slot of a forward entry (x, y) of anchor i from packed bytes. Each bit of the slot is [P > Q]
// for small counts 0 <= P, Q <= 64, read as bit 7 of the byte 127 + P - Q (in 63..191, so four
// such bytes add and subtract in one 32-bit word without carries between them):
//   word A, bytes 0-3: x only   [xfree + g > C_xy]     word B: xy only [C_xy > g]
//                      y only   [yfree + g > C_xy]             yi only [C_yi > g]
//                      i only   [ifree + g > C_yi]             xi only [C_xi > g]
//                      b        [ifree > C_yi]                 b2      [yfree > C_xy]
// with xfree = s_x - C_xi, ifree = s_i - C_xi, yfree = s_y - C_yi and g = |x ∩ y ∩ i|. KA and
// KB hold x's constants (bytes 127 + xfree, 127, ifree, ifree and 127, 127, 127 + C_xi, 127), y
// comes in as W16. Two byte dot products with the slot bit weights gather the bits (times 128,
// plus `base`); with g = 0 the slot is kTSlotNoTriple + low when y overlaps i and
// kTSlotOpenX + (X | Y << 1 | I << 2) when it does not, which are the x-only, y-only and i-only
// bits then (C_yi = 0), told apart by the yi-only bit. Returns 128 slot + base: the byte
// offset of counter (slot, copy) of 32 copies per slot, copy fastest, for base = 4 copy.
__device__ inline uint32_t tslot_offset(uint32_t KA, uint32_t KB, uint32_t w16, uint32_t cxy,
                                        uint32_t g, uint32_t base) {
  const uint32_t gm = g * 0x10101u;
  const uint32_t ya = __builtin_amdgcn_perm(0u, w16, 0x0000010cu) ^ 0x7F7F0000u;  // 0, yfree, 127 - C_yi, 127 - C_yi
  const uint32_t yb = __builtin_amdgcn_perm(0u, w16, 0x010c000cu);                // 0, C_yi, 0, yfree
  const uint32_t c24 = __builtin_amdgcn_perm(0u, cxy, 0x000c0c0cu);               // C_xy << 24
  const uint32_t A = (KA + ya + gm - cxy * 0x101u) & 0x80808080u;
  const uint32_t B = (KB + yb + cxy - gm - c24) & 0x80808080u;
  // slot bits: A -> 2, 3, 4, 1; B -> 5, 6, 7, 0 (times 128 by the byte value 128 of a set bit)
  const uint32_t s1 = __builtin_amdgcn_udot4(
      B, 32u | 64u << 8 | 128u << 16 | 1u << 24,
      __builtin_amdgcn_udot4(A, 4u | 8u << 8 | 16u << 16 | 2u << 24, base, false), false);
  // g = 0: low bits 0-2, the yi-only bit 3, from slot 256
  const uint32_t s0 = __builtin_amdgcn_udot4(
      B, 8u << 8, __builtin_amdgcn_udot4(A, 1u | 2u << 8 | 4u << 16, base + 128u * 256u, false),
      false);
  return g ? s1 : s0;
}

// Counts, for every ordered pair (x earlier, y later) of the nN neighbours of the anchor (in time
// order), the as-if-open id (8 cases: fx, fy, [C_xi + C_yi < si]) with one wave. load(k, C, f)
// gives neighbour k's overlap C with the anchor and f = [it has a node outside the anchor]. The
// running class histogram of earlier entries is kept across lanes (lane c holds the number of
// earlier entries with C = c + 1), and earlier entries of the same 64-entry chunk are counted
// with one ballot per (f, C) class.
template <typename Load>
__device__ inline void ordered_open_pairs(uint32_t nN, int si, Load&& load, unsigned long long cnt[8]) {
  const int lane = lane_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned long long R0 = 0, R1 = 0;  // earlier entries with f = 0 / 1 and C == lane + 1
  for (int k = 0; k < 8; ++k) cnt[k] = 0;
  for (uint32_t c0 = 0; c0 < nN; c0 += kWave) {
    const uint32_t k = c0 + lane;
    const bool valid = k < nN;
    int C = 0, f = 0;
    if (valid) load(k, C, f);
    const int T = si - C;  // earlier x count towards [C_x + C < si] iff C_x < T
    // earlier chunks: prefix over the lane-distributed histograms
    unsigned long long P0 = R0, P1 = R1;  // inclusive prefix over lanes 0..lane
    for (int o = 1; o < kWave; o <<= 1) {
      const unsigned long long q0 = __shfl_up(P0, o), q1 = __shfl_up(P1, o);
      if (lane >= o) { P0 += q0; P1 += q1; }
    }
    const unsigned long long tot0 = __shfl(P0, kWave - 1), tot1 = __shfl(P1, kWave - 1);
    // earlier entries with C_x < T, i.e. histogram lanes 0 .. T - 2
    const int src = T >= 2 ? T - 2 : 0;
    unsigned long long a0 = __shfl(P0, src), a1 = __shfl(P1, src);
    if (T < 2) { a0 = 0; a1 = 0; }
    unsigned long long n0 = tot0, n1 = tot1;
    // same chunk
    const unsigned long long all0 = __ballot(valid && f == 0), all1 = __ballot(valid && f == 1);
    n0 += (unsigned long long)__popcll(all0 & below);
    n1 += (unsigned long long)__popcll(all1 & below);
    for (int c = 1; c <= si; ++c) {
      const unsigned long long m0 = __ballot(valid && f == 0 && C == c);
      const unsigned long long m1 = __ballot(valid && f == 1 && C == c);
      if (c < T) {
        a0 += (unsigned long long)__popcll(m0 & below);
        a1 += (unsigned long long)__popcll(m1 & below);
      }
      if (lane == c - 1) {
        R0 += (unsigned long long)__popcll(m0);
        R1 += (unsigned long long)__popcll(m1);
      }
    }
    if (valid) {
      // id bits: fx (x earlier) + 2 fy (this entry) + 4 [C_x + C < si]
      cnt[0 + 2 * f + 4] += a0;
      cnt[0 + 2 * f] += n0 - a0;
      cnt[1 + 2 * f + 4] += a1;
      cnt[1 + 2 * f] += n1 - a1;
    }
  }
  for (int k = 0; k < 8; ++k) cnt[k] = wave_sum(cnt[k]);
}

// Histogram: 32 copies of the kTSlots 32-bit counters shared by the workgroup (copy = lane & 31,
// copies of a slot adjacent), updated with per-lane LDS atomics; a counter holds counts of one
// anchor (the workgroup adds at most W^2 per anchor, W <= ~19k window entries from the LDS
// limit, so < 2^32). About half of the temporal workload's closed triangles share one Venn id,
// so the copies decide the same-address conflicts: 2 copies per wave 613 ms, 16 shared copies
// 415 ms, 32 shared copies 408 ms (34.8 KB, as for 2 per wave: two workgroups per CU).
// The neighbours of the anchor are its backward row B(i) (ascending, with C_xi).
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_tcount(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vpos, const int32_t* __restrict__ lo,
    const unsigned long long* __restrict__ foff, const unsigned long long* __restrict__ fdeg,
    const uint32_t* __restrict__ frows, const unsigned long long* __restrict__ boff,
    const unsigned long long* __restrict__ bdeg, const uint32_t* __restrict__ brows,
    int32_t Wmax, unsigned long long* __restrict__ acc) {
  constexpr int kC = 32;  // counter copies per slot, shared by the workgroup (copy = lane & 31)
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned int hist[kTSlots * kC];
  __shared__ unsigned long long bsum[kTemporalBins];
  __shared__ int8_t tbl[128];
  __shared__ uint32_t s_next;
  // anchors in time order: the workgroups in flight share most of their windows, so the rows
  // they walk are re-read from the Infinity Cache (most work first instead: 412 vs 393 ms and
  // 2.64 vs 1.96 TB fetched past L2 per launch on the 200k-edge temporal workload)
  const int32_t i = (int32_t)blockIdx.x;
  const int32_t base = lo[i];
  const int32_t W = i - base;
  const uint32_t nN = (uint32_t)bdeg[i];
  if (W < 2 || nN == 0) return;
  const size_t off = ((size_t)Wmax * sizeof(MT) + 15) & ~(size_t)15;
  MT* M = reinterpret_cast<MT*>(smem);
  uint16_t* W16 = reinterpret_cast<uint16_t*>(smem + off);

  for (int k = threadIdx.x; k < kTSlots * kC; k += blockDim.x) hist[k] = 0;
  for (int k = threadIdx.x; k < kTemporalBins; k += blockDim.x) bsum[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_temporal_bin[threadIdx.x];
  if (threadIdx.x == 0) s_next = 0;
  build_window_masks<MT>(M, i, base, i - 1, eoff, enodes, voff, vpos);
  for (int k = threadIdx.x; k < W; k += blockDim.x) {
    const uint32_t c = (uint32_t)popc(M[k]);
    W16[k] = (uint16_t)(c | ((uint32_t)esize[base + k] - c) << 8);
  }
  __syncthreads();

  const int lane = lane_id(), w = wave_id();
  const int si = (int)esize[i];
  const uint32_t* NB = brows + boff[i];
  unsigned char* hwb = reinterpret_cast<unsigned char*>(hist);
  const uint32_t cbase = 4u * (uint32_t)(lane & (kC - 1));
  // the window tables addressed by absolute position (the window start folded into the base)
  const MT* Mabs = M - base;
  const uint16_t* Wabs = W16 - base;

  // open at i, counted for all ordered neighbour pairs as if disjoint (closed ones are taken
  // back below when their triangle is met)
  if (w == NW - 1) {
    unsigned long long cnt[8];
    ordered_open_pairs(nN, si, [&](uint32_t k, int& C, int& f) {
      const uint32_t e = NB[k];
      C = (int)(e & kCntMask);
      f = (Wabs[e >> kCntBits] >> 8) != 0 ? 1 : 0;
    }, cnt);
    if (lane == 0)
      for (int k = 0; k < 8; ++k)
        if (cnt[k]) atomicAdd(&acc[kTAccHist + tbl[48 + k]], cnt[k]);
  }

  // For every neighbour x (one wave each) the forward row of x inside (x, i): every entry y is a
  // closed triangle (x, y, i) when y overlaps i, else a triple open at x with y between x and i;
  // both kinds are one LDS atomic into the slot tslot_offset gives. The triples open at x with y
  // before x (side (e)) are counted by k_twedge for every pair (y in B(x), i in F(x)) with y
  // inside i's window, whether or not y overlaps i; the pairs that do are the triangles (y, x, i),
  // met here as the forward entries of their earliest edge, and each takes its pair back through
  // its slot.
  constexpr int kTLoads = 8;  // forward-row loads in flight per lane
  // entry e = y << kCntBits | C_xy (C_xy >= 1) is before the anchor iff e - 1 < lim; a read past
  // the row returns 0, for which e - 1 wraps past every lim
  const uint32_t lim = ((uint32_t)i << kCntBits) - 1u;
  for (;;) {
    uint32_t xi = 0;
    if (lane == 0) xi = atomicAdd(&s_next, 1u);
    xi = __shfl(xi, 0);
    if (xi >= nN) break;
    const uint32_t ex = NB[xi];
    const int32_t x = (int32_t)(ex >> kCntBits);
    const uint32_t cxi = ex & kCntMask;
    const MT mx = Mabs[x];
    const uint32_t xfree = (uint32_t)Wabs[x] >> 8;
    const uint32_t ifree = (uint32_t)si - cxi;
    const uint32_t KA = (xfree + 127u) | 127u << 8 | ifree << 16 | ifree << 24;
    const uint32_t KB = 127u | 127u << 8 | (127u + cxi) << 16 | 127u << 24;
    // Rows are read through a buffer descriptor bounded by the row's length, and the scan stops
    // at the first chunk that reaches the anchor (rows are ascending).
    const uint32_t n = (uint32_t)fdeg[x];
    const __amdgpu_buffer_rsrc_t rs = row_rsrc(frows + foff[x], n);
    auto visit = [&](uint32_t e) {
      if (e - 1u >= lim) return;
      const uint32_t y = e >> kCntBits;
      const uint32_t g = (uint32_t)popc((MT)(mx & Mabs[y]));
      atomicAdd(reinterpret_cast<unsigned int*>(hwb + tslot_offset(KA, KB, Wabs[y], e & kCntMask, g, cbase)), 1u);
    };
    for (uint32_t j0 = 0; j0 < n; j0 += kTLoads * kWave) {
      uint32_t e[kTLoads];
#pragma unroll
      for (int u = 0; u < kTLoads; ++u) e[u] = row_load(rs, j0 + lane + u * kWave);
#pragma unroll
      for (int u = 0; u < kTLoads; ++u) visit(e[u]);
      if (__ballot(e[kTLoads - 1] - 1u >= lim) != 0ull) break;
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kTSlots; p += blockDim.x) {
    unsigned long long t = 0;
    for (int c = 0; c < kC; ++c) t += hist[p * kC + c];
    if (t) tslot_fold(p, t, bsum, tbl);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kTemporalBins; b += blockDim.x)
    if (bsum[b]) atomicAdd(&acc[kTAccHist + b], bsum[b]);
}

// Side (e) of the triples open at a middle edge x: for every pair (y in B(x), i in F(x)) with y
// inside i's window (lo_i <= y; then also lo_i <= x), the bin of (y, x, i) as if y and i were
// disjoint — Venn id Y | X << 1 | I << 2 | 8 | 16 with Y = [y ⊄ x], X = [|x| > C_xy + C_xi],
// I = [i ⊄ x]. k_tcount takes back the pairs with y overlapping i (the triangles (y, x, i)).
// One wave per x walks F(x) in time order, 64 anchors i per step (one per lane; lo_i is
// non-decreasing along F(x)), and keeps the entries of B(x) inside the step's first window in an
// LDS histogram by (Y, C_xy) with its prefix sum over C_xy across the lanes. A lane's count is the
// histogram's "C_xy < |x| - C_xi" minus the entries leaving the window during the step that lie
// before its own window start (those are broadcast one by one).
template <int NW>
__global__ __launch_bounds__(NW * kWave) void k_twedge(
    int32_t E, const int32_t* __restrict__ lo, const uint8_t* __restrict__ esize,
    const unsigned long long* __restrict__ foff, const unsigned long long* __restrict__ fdeg,
    const uint32_t* __restrict__ frows, const unsigned long long* __restrict__ boff,
    const unsigned long long* __restrict__ bdeg, const uint32_t* __restrict__ brows,
    unsigned long long* __restrict__ acc) {
  __shared__ uint32_t H[NW][2][kWave + 1];  // [Y][C_xy], C_xy in 1..64
  __shared__ int8_t tbl[128];
  const int lane = lane_id(), w = wave_id();
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_temporal_bin[threadIdx.x];
  for (int k = lane; k < 2 * (kWave + 1); k += kWave) (&H[w][0][0])[k] = 0u;
  __syncthreads();
  const int32_t x = (int32_t)blockIdx.x * NW + w;
  if (x >= E) return;
  const uint32_t nb = (uint32_t)bdeg[x], nf = (uint32_t)fdeg[x];
  if (nb == 0 || nf == 0) return;
  const int sx = esize[x];
  const uint32_t* B = brows + boff[x];
  const uint32_t* F = frows + foff[x];
  auto ybit = [&](uint32_t e) { return (int)esize[e >> kCntBits] > (int)(e & kCntMask) ? 1 : 0; };
  for (uint32_t j = lane; j < nb; j += kWave) {
    const uint32_t e = B[j];
    atomicAdd(&H[w][ybit(e)][e & kCntMask], 1u);
  }
  __threadfence_block();
  // inclusive prefix over C_xy of the histogram (lane c - 1 holds the count of C_xy <= c)
  uint32_t P0 = 0, P1 = 0;
  auto rescan = [&]() {
    P0 = H[w][0][lane + 1];
    P1 = H[w][1][lane + 1];
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t q0 = __shfl_up(P0, o), q1 = __shfl_up(P1, o);
      if (lane >= o) {
        P0 += q0;
        P1 += q1;
      }
    }
  };
  rescan();
  // entries of B(x) from p on are inside the current window; chunk [cb, cb + 64) in registers
  uint32_t p = 0, cb = 0;
  uint32_t eb = lane < nb ? B[lane] : 0u;
  int yb = lane < nb ? ybit(eb) : 0;
  unsigned long long cnt[2][2][2] = {};  // per lane: [I][X][Y]
  for (uint32_t j0 = 0; j0 < nf; j0 += kWave) {
    const uint32_t j = j0 + lane;
    uint32_t ei = 0;
    int32_t loi = 0x7FFFFFFF;
    int si = 0;
    if (j < nf) {
      ei = F[j];
      loi = lo[ei >> kCntBits];
      si = esize[ei >> kCntBits];
    }
    const bool valid = loi <= x;  // a prefix of the lanes (lo_i non-decreasing)
    const unsigned long long vb = __ballot(valid);
    if (vb == 0ull) break;        // x is outside every later anchor's window
    const int32_t lo_last = __builtin_amdgcn_readlane(loi, __popcll(vb) - 1);
    const int cxi = (int)(ei & kCntMask);
    const int T = sx - cxi;  // X = [C_xy < T]
    // counts over the window of the step's first anchor (histogram as it stands)
    uint32_t a0 = __shfl(P0, T >= 2 ? T - 2 : 0), a1 = __shfl(P1, T >= 2 ? T - 2 : 0);
    if (T < 2) {
      a0 = 0u;
      a1 = 0u;
    }
    uint32_t t0 = __builtin_amdgcn_readlane(P0, kWave - 1), t1 = __builtin_amdgcn_readlane(P1, kWave - 1);
    // entries leaving the window during this step (y < lo of the step's last anchor): each lane
    // takes back those before its own window start, then they leave the histogram
    bool dropped = false;
    while (p < nb) {
      const uint32_t idx = cb + (uint32_t)lane;
      const bool rm = idx >= p && idx < nb && (int32_t)(eb >> kCntBits) < lo_last;
      unsigned long long bal = __ballot(rm);
      const uint32_t nrm = (uint32_t)__popcll(bal);
      while (bal) {
        const int sl = __builtin_ctzll(bal);
        bal &= bal - 1ull;
        const uint32_t ee = __builtin_amdgcn_readlane(eb, sl);
        const int ye = __builtin_amdgcn_readlane(yb, sl);
        if ((int32_t)(ee >> kCntBits) < loi) {
          const uint32_t lt = (int)(ee & kCntMask) < T ? 1u : 0u;
          if (ye) {
            t1 -= 1u;
            a1 -= lt;
          } else {
            t0 -= 1u;
            a0 -= lt;
          }
        }
      }
      if (rm) atomicSub(&H[w][yb][eb & kCntMask], 1u);
      dropped = dropped || nrm != 0;
      p += nrm;
      if (p < cb + kWave) break;  // the chunk still holds entries inside the window
      cb += kWave;                // chunk used up: fetch the next one
      eb = cb + lane < nb ? B[cb + lane] : 0u;
      yb = cb + lane < nb ? ybit(eb) : 0;
    }
    if (valid) {
      const int I = si > cxi ? 1 : 0;
      cnt[I][1][0] += a0;
      cnt[I][0][0] += t0 - a0;
      cnt[I][1][1] += a1;
      cnt[I][0][1] += t1 - a1;
    }
    if (dropped) {
      __threadfence_block();
      rescan();
    }
    if (vb != ~0ull) break;  // some anchor of this step already lies past x's reach
  }
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int X = 0; X < 2; ++X)
#pragma unroll
      for (int Y = 0; Y < 2; ++Y) {
        const unsigned long long c = wave_sum(cnt[I][X][Y]);
        if (lane == 0 && c) atomicAdd(&acc[kTAccHist + tbl[Y | X << 1 | I << 2 | 8 | 16]], c);
      }
}

// ---------------------------------------------------------------------------------------------
// Tiled kernels: windows longer than one LDS table. Positions are cut into tiles of T positions
// (relative to the window start); a table is built per tile from the part of each node's position
// list inside it (tile_ranges / for_tile_entries, positions ascending per node).
// ---------------------------------------------------------------------------------------------

// fdeg / bdeg of x over position tiles of [lo_x, hi_x].
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_tproj_count_tiled(
    int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vpos,
    const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    unsigned long long* __restrict__ fdeg, unsigned long long* __restrict__ bdeg) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ TileRanges tr;
  __shared__ uint32_t red[2][kWavesPerBlock];
  const int32_t x = blockIdx.x;
  const uint32_t s = eoff[x];
  const int sx = (int)(eoff[x + 1] - s);
  for (int32_t k = threadIdx.x; k < T; k += blockDim.x) M[k] = (MT)0;
  uint32_t f = 0, b = 0;
  for (int32_t t0 = lo[x]; t0 <= hi[x]; t0 += T) {
    const int32_t t1 = min(hi[x] + 1, t0 + T);
    const uint32_t total = tile_ranges(tr, sx, enodes + s, voff, vpos, t0, t1);
    if (total == 0) continue;
    for_tile_entries(tr, sx, total, vpos, [&](int k, int32_t y) {
      if (y != x) atomicOr(&M[y - t0], (MT)1 << k);
    });
    __syncthreads();
    for_tile_entries(tr, sx, total, vpos, [&](int k, int32_t y) {
      if (y == x) return;
      const MT m = M[y - t0];
      if (m && low_bit(m) == k) {
        if (y > x) ++f;
        else ++b;
        M[y - t0] = (MT)0;
      }
    });
  }
  f = wave_sum(f);
  b = wave_sum(b);
  if (lane_id() == 0) {
    red[0][wave_id()] = f;
    red[1][wave_id()] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ff = 0, bb = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
      ff += red[0][w];
      bb += red[1][w];
    }
    fdeg[x] = ff;
    bdeg[x] = bb;
  }
}

// Compacts positions y in [from, to] with M[y - base] != 0 into out[] (ascending), and their masks
// into outm[] when it is not null. Returns the number written (uniform).
template <typename MT>
__device__ inline uint32_t compact_rows_masked(const MT* M, int32_t base, int32_t from, int32_t to,
                                               uint32_t* __restrict__ out, MT* __restrict__ outm,
                                               uint32_t* wcnt) {
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t running = 0;
  for (int32_t y0 = from; y0 <= to; y0 += blockDim.x) {
    const int32_t y = y0 + (int32_t)threadIdx.x;
    const MT m = (y <= to) ? M[y - base] : (MT)0;
    const unsigned long long bal = __ballot(m != (MT)0);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) {
      woff += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    if (m) {
      const uint32_t pos = running + woff + (uint32_t)__popcll(bal & below);
      out[pos] = ((uint32_t)y << kCntBits) | (uint32_t)popc(m);
      if (outm) outm[pos] = m;
    }
    running += tot;
    __syncthreads();
  }
  return running;
}

// F(x) and B(x) over position tiles (ascending), and bmask = x's nodes contained in each B(x)
// entry (the masks the tiled counting kernel needs for the anchor's neighbours).
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_tproj_fill_tiled(
    int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vpos,
    const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    const unsigned long long* __restrict__ foff, const unsigned long long* __restrict__ boff,
    uint32_t* __restrict__ frows, uint32_t* __restrict__ brows, MT* __restrict__ bmask) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ TileRanges tr;
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t x = blockIdx.x;
  const uint32_t s = eoff[x];
  const int sx = (int)(eoff[x + 1] - s);
  for (int32_t k = threadIdx.x; k < T; k += blockDim.x) M[k] = (MT)0;
  uint32_t* F = frows + foff[x];
  uint32_t* B = brows + boff[x];
  MT* BM = bmask + boff[x];
  uint32_t nf = 0, nb = 0;
  for (int32_t t0 = lo[x]; t0 <= hi[x]; t0 += T) {
    const int32_t t1 = min(hi[x] + 1, t0 + T);
    const uint32_t total = tile_ranges(tr, sx, enodes + s, voff, vpos, t0, t1);
    if (total == 0) continue;
    for_tile_entries(tr, sx, total, vpos, [&](int k, int32_t y) {
      if (y != x) atomicOr(&M[y - t0], (MT)1 << k);
    });
    __syncthreads();
    nb += compact_rows_masked<MT>(M, t0, t0, min(t1, x) - 1, B + nb, BM + nb, wcnt);
    nf += compact_rows_masked<MT>(M, t0, max(t0, x + 1), t1 - 1, F + nf, (MT*)nullptr, wcnt);
    for_tile_entries(tr, sx, total, vpos, [&](int, int32_t y) {
      if (y != x) M[y - t0] = (MT)0;
    });
  }
}

// Counting over position tiles of the window [lo_i, i). The anchor's neighbours are its
// backward row B(i) (positions ascending, with C_xi) and their masks bmask; the open-at-anchor
// pairs come from them directly. For each tile a table of the anchor's masks over the tile's
// positions answers "is y a neighbour of i" for the entries of F(x) inside the tile; per (i, x)
// read positions (fcur, aligned with B(i)) carry each row from tile to tile. The triples open at
// x with the other edge earlier come from k_twedge, as for the one-table kernel.
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_tcount_tiled(
    int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vpos, const int32_t* __restrict__ lo,
    const unsigned long long* __restrict__ foff, const unsigned long long* __restrict__ fdeg,
    const uint32_t* __restrict__ frows, const unsigned long long* __restrict__ boff,
    const unsigned long long* __restrict__ bdeg, const uint32_t* __restrict__ brows,
    const MT* __restrict__ bmask, uint32_t* __restrict__ fcur,
    const int32_t* __restrict__ order, unsigned long long* __restrict__ acc) {
  constexpr int kC = 32;  // counter copies per slot, shared by the workgroup, as in k_tcount
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ unsigned int hist[kTSlots * kC];
  __shared__ unsigned long long bsum[kTemporalBins];
  __shared__ int8_t tbl[128];
  __shared__ TileRanges tr;
  const int32_t i = order[blockIdx.x];
  const int32_t base = lo[i];
  const uint32_t nN = (uint32_t)bdeg[i];
  if (i - base < 2 || nN == 0) return;
  for (int k = threadIdx.x; k < kTSlots * kC; k += blockDim.x) hist[k] = 0;
  for (int k = threadIdx.x; k < kTemporalBins; k += blockDim.x) bsum[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_temporal_bin[threadIdx.x];
  for (int32_t k = threadIdx.x; k < T; k += blockDim.x) M[k] = (MT)0;
  const uint32_t s = eoff[i];
  const int si = (int)(eoff[i + 1] - s);
  const uint32_t* NL = brows + boff[i];
  const MT* NM = bmask + boff[i];
  uint32_t* FC = fcur + boff[i];
  const int lane = lane_id(), w = wave_id();
  __syncthreads();  // tbl before the open pairs

  // open at i, counted for all ordered neighbour pairs as if disjoint
  if (w == NW - 1) {
    unsigned long long cnt[8];
    ordered_open_pairs(nN, si, [&](uint32_t k, int& C, int& f) {
      const uint32_t e = NL[k];
      C = (int)(e & kCntMask);
      f = (int)esize[e >> kCntBits] > C ? 1 : 0;
    }, cnt);
    if (lane == 0)
      for (int k = 0; k < 8; ++k)
        if (cnt[k]) atomicAdd(&acc[kTAccHist + tbl[48 + k]], cnt[k]);
  }

  unsigned int* hw = hist + (lane & (kC - 1));
  // triples open at x with the other edge later, per lane by (I, X + 2Y) as in k_tcount (those
  // with the other edge earlier: k_twedge, less the triangles taken back through their slots)
  uint32_t open_cnt[2][4] = {};
  for (int32_t t0 = base; t0 < i; t0 += T) {
    const int32_t t1 = min(i, t0 + T);
    const uint32_t total = tile_ranges(tr, si, enodes + s, voff, vpos, t0, t1);
    for_tile_entries(tr, si, total, vpos, [&](int k, int32_t y) { atomicOr(&M[y - t0], (MT)1 << k); });
    __syncthreads();
    for (uint32_t xi = (uint32_t)w; xi < nN; xi += NW) {
      const uint32_t ex = NL[xi];
      const int32_t x = (int32_t)(ex >> kCntBits);
      const int cxi = (int)(ex & kCntMask);
      const MT mx = NM[xi];
      const int sx = esize[x];
      const int xfree = sx - cxi, ifree = si - cxi;
      unsigned long long fields_b = 0;
      // (b) forward row of x inside [max(x + 1, t0), t1): closed triangles and triples open at x
      // (the read position starts at 0 in the first tile, also for an x beyond that tile)
      uint32_t fpos = (t0 == base) ? 0u : FC[xi];
      if (x + 1 < t1) {
        const uint32_t* row = frows + foff[x];
        const uint32_t n = (uint32_t)fdeg[x];
        uint32_t& pos = fpos;
        for (;;) {
          const uint32_t j = pos + lane;
          bool v = false;
          uint32_t e = 0;
          if (j < n) {
            e = row[j];
            v = (int32_t)(e >> kCntBits) < t1;
          }
          if (v) {
            const int32_t y = (int32_t)(e >> kCntBits);
            const int cxy = (int)(e & kCntMask);
            const MT my = M[y - t0];
            const int sy = esize[y];
            if (my) {
              const int cyi = popc(my);
              const int g = popc((MT)(mx & my));
              const unsigned low = nz(xfree - cxy + g) | nz(sy - cxy - cyi + g) << 1 | nz(ifree - cyi + g) << 2;
              const unsigned mid = nz(cxy - g) | nz(cyi - g) << 1 | nz(cxi - g) << 2;
              const unsigned b = (unsigned)max(min(ifree - cyi, 1), 0);
              const unsigned b2 = (unsigned)max(min(sy - cyi - cxy, 1), 0);
              const unsigned slot = g ? (((low << 1 | b) + (mid << 4)) << 1 | b2) : ((unsigned)kTSlotNoTriple | low);
              atomicAdd(&hw[slot * kC], 1u);
            } else {
              fields_b += 1ull << (nz(xfree - cxy) << 4 | nz(sy - cxy) << 5);
            }
          }
          const uint32_t nv = (uint32_t)__popcll(__ballot(v));
          pos += nv;
          if (nv < (uint32_t)kWave) break;
        }
      }
      if (lane == 0) FC[xi] = fpos;
      // (I is uniform in the wave; both branches index the counters with constants)
      if (si > cxi) {
#pragma unroll
        for (int k = 0; k < 4; ++k) open_cnt[1][k] += (uint32_t)(fields_b >> (16 * k)) & 0xFFFFu;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) open_cnt[0][k] += (uint32_t)(fields_b >> (16 * k)) & 0xFFFFu;
      }
    }
    __syncthreads();
    // this tile's triangle slots into their closed (+) and as-if-open (-) bins (per tile, so a
    // slot counter stays far below 2^32)
    for (int p = threadIdx.x; p < kTSlots; p += blockDim.x) {
      unsigned long long sum = 0;
      for (int c = 0; c < kC; ++c) {
        sum += hist[p * kC + c];
        hist[p * kC + c] = 0;
      }
      if (sum) tslot_fold(p, sum, bsum, tbl);
    }
    for_tile_entries(tr, si, total, vpos, [&](int, int32_t y) { M[y - t0] = (MT)0; });
  }
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t c = wave_sum(open_cnt[I][k]);
      const int X = k & 1, Y = k >> 1;
      if (lane == 0 && c) atomicAdd(&bsum[tbl[X | Y << 1 | I << 2 | 8 | 32]], (unsigned long long)c);
    }
  __syncthreads();
  for (int b = threadIdx.x; b < kTemporalBins; b += blockDim.x)
    if (bsum[b]) atomicAdd(&acc[kTAccHist + b], bsum[b]);
}

// Per-anchor work estimate (for launch order) and statistics.
__global__ void k_tstats(int32_t E, const int32_t* __restrict__ lo,
                         const unsigned long long* __restrict__ fdeg,
                         const unsigned long long* __restrict__ bdeg,
                         unsigned long long* __restrict__ work, unsigned long long* __restrict__ acc) {
  unsigned long long pairs = 0, mxw = 0, np = 0;
  for (int32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < E; x += gridDim.x * blockDim.x) {
    const unsigned long long b = bdeg[x];
    pairs += fdeg[x];
    mxw = max(mxw, (unsigned long long)(x - lo[x]));
    np += b * (b - (b > 0)) / 2;
    work[x] = b * b + b;
  }
  pairs = wave_sum(pairs);
  np = wave_sum(np);
  mxw = wave_max(mxw);
  if (lane_id() == 0) {
    atomicAdd(&acc[kTAccPairs], pairs);
    atomicAdd(&acc[kTAccNbrPairs], np);
    atomicMax(&acc[kTAccMaxWin], mxw);
  }
}

// sum over forward pairs (x, y), y > x, of the window degree of y (byte model of the count kernel)
__global__ __launch_bounds__(kBlock) void k_torient(int32_t E, const unsigned long long* __restrict__ foff,
                                                    const unsigned long long* __restrict__ fdeg,
                                                    const unsigned long long* __restrict__ bdeg,
                                                    const uint32_t* __restrict__ frows,
                                                    unsigned long long* __restrict__ acc) {
  const int32_t x = blockIdx.x * kWavesPerBlock + wave_id();
  unsigned long long s = 0;
  if (x < E) {
    const uint32_t* row = frows + foff[x];
    const uint32_t n = (uint32_t)fdeg[x];
    for (uint32_t j = lane_id(); j < n; j += kWave) {
      const int32_t y = (int32_t)(row[j] >> kCntBits);
      s += fdeg[y] + bdeg[y];
    }
  }
  s = wave_sum(s);
  if (lane_id() == 0 && s) atomicAdd(&acc[kTAccOrient], s);
}

__global__ void k_tzero(unsigned long long* p, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}
__global__ void k_tiota(int32_t* p, int32_t n) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = k;
}

}  // namespace hm

// ---------------------------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------------------------
struct hm_tgraph {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t E = 0, V = 0, max_size = 0;
  int64_t I = 0;
  hm::DevBuf eoff, enodes, esize, voff, vpos, times;
  // scratch
  hm::DevBuf lo, hi, fdeg, bdeg, foff, boff, frows, brows, work, work2, ids, order, acc, tmp;
  hm::DevBuf bmask, fcur;  // tiled windows: anchor masks and per-(anchor, neighbour) read positions
  int32_t tile = 0;              // 0 = automatic, > 0 = forced position tiles
  int32_t last_tile = 0;         // tile of the last count (0 = one table per window)
  std::vector<int32_t> order_to_input;  // position -> input index
  hm_temporal_stats stats{};
  bool have_stats = false;
  float ms_project = 0.f, ms_count = 0.f, ms_wedge = 0.f;
};

namespace hm {

static size_t win_lds_bytes(int32_t R, bool wide) {
  return (((size_t)R * (wide ? 8 : 4)) + 15) & ~(size_t)15;
}

// Launch of the counting kernel: 16 waves per anchor workgroup and per-lane 64-bit LDS atomics
// for the histogram. Measured on the 200k-edge temporal workload (c4): 4 waves with wave-aggregated
// histograms 7.04 s, 8 waves aggregated 3.74 s, 8 waves atomics 2.46 s, 16 waves atomics 2.12 s.
constexpr int kTCountWaves = 16;
constexpr int kWedgeWaves = 4;  // k_twedge: one wave per middle edge
// Static LDS of the tiled kernels (tile walk ranges, reductions; for counting also the 34.8 KB
// triangle-slot counters, the bin sums and the bin table), kept free next to the position table.
constexpr size_t kTProjTiledStaticLds = 4 * 1024;
constexpr size_t kTCountTiledStaticLds = 36 * 1024;

template <typename MT>
static hm_status launch_tcount(hm_tgraph* g, int32_t Wmax, size_t lds, unsigned long long* acc) {
  const int32_t E = g->E;
  HM_TRY(allow_lds((const void*)k_tcount<MT, kTCountWaves>, lds));
  k_tcount<MT, kTCountWaves><<<E, kTCountWaves * kWave, lds, g->stream>>>(
      E, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->esize.as<uint8_t>(),
      g->voff.as<uint32_t>(), g->vpos.as<int32_t>(), g->lo.as<int32_t>(),
      g->foff.as<unsigned long long>(), g->fdeg.as<unsigned long long>(), g->frows.as<uint32_t>(),
      g->boff.as<unsigned long long>(), g->bdeg.as<unsigned long long>(), g->brows.as<uint32_t>(),
      Wmax, acc);
  HM_HIP(hipGetLastError());
  return HM_OK;
}

template <typename MT>
static hm_status count_temporal_impl(hm_tgraph* g, long long delta, unsigned long long* h_acc) {
  const int32_t E = g->E;
  hipStream_t st = g->stream;
  const unsigned eb = (unsigned)((E + kBlock - 1) / kBlock);
  HM_TRY(g->lo.ensure(sizeof(int32_t) * E));
  HM_TRY(g->hi.ensure(sizeof(int32_t) * E));
  HM_TRY(g->fdeg.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(g->bdeg.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(g->foff.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(g->boff.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(g->work.ensure(sizeof(unsigned long long) * E));
  HM_TRY(g->work2.ensure(sizeof(unsigned long long) * E));
  HM_TRY(g->ids.ensure(sizeof(int32_t) * E));
  HM_TRY(g->order.ensure(sizeof(int32_t) * E));
  HM_TRY(g->acc.ensure(sizeof(unsigned long long) * kTAccCount));
  unsigned long long* acc = g->acc.as<unsigned long long>();

  HM_HIP(hipEventRecord(g->ev[0], st));
  k_tzero<<<1, 128, 0, st>>>(acc, kTAccCount);
  k_twindow<<<eb, kBlock, 0, st>>>(E, g->times.as<long long>(), delta, g->lo.as<int32_t>(),
                                   g->hi.as<int32_t>());
  // widest projection range decides the LDS of the projection kernels
  std::vector<int32_t> hlo((size_t)E), hhi((size_t)E);
  HM_HIP(hipMemcpyAsync(hlo.data(), g->lo.p, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(hhi.data(), g->hi.p, sizeof(int32_t) * E, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  int32_t Rmax = 1, Wmax = 1;
  for (int32_t x = 0; x < E; ++x) {
    Rmax = std::max(Rmax, hhi[(size_t)x] - hlo[(size_t)x] + 1);
    Wmax = std::max(Wmax, x - hlo[(size_t)x]);
  }
  int lim = 65536;
  (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device);
  const size_t lds_proj = win_lds_bytes(Rmax, sizeof(MT) == 8);
  const size_t lds_cnt = TCountLds<MT>::bytes(Wmax);
  const size_t static_cnt = sizeof(uint32_t) * 16 * kTSlots * 2 +
                            sizeof(unsigned long long) * kTemporalBins + 256;
  // one LDS table per window when it fits; position tiles otherwise (or when forced)
  const bool tiled = g->tile > 0 || lds_proj > (size_t)lim || lds_cnt + static_cnt > (size_t)lim;
  int32_t Tp = 0, Tc = 0;
  if (tiled) {
    const int32_t most_p = (int32_t)((((size_t)lim - kTProjTiledStaticLds) / sizeof(MT)) & ~(size_t)63);
    const int32_t most_c = (int32_t)((((size_t)lim - kTCountTiledStaticLds) / sizeof(MT)) & ~(size_t)63);
    Tp = g->tile > 0 ? g->tile : most_p;
    Tc = g->tile > 0 ? g->tile : most_c;
    if (Tp > most_p || Tc > most_c) {
      set_error("temporal count: a tile of %d positions exceeds the LDS limit", g->tile);
      return HM_ERR_BAD_INPUT;
    }
    // k_tcount_tiled's slot counters are 32-bit and flushed per tile: within one tile an anchor
    // meets at most Wmax window neighbours, each with at most Tc forward entries in the tile
    if ((unsigned long long)Wmax * (unsigned long long)Tc >= (1ull << 32)) {
      set_error("temporal count: window of %d x tile of %d positions exceeds the bound of the "
                "32-bit slot counters (2^32)", Wmax, Tc);
      return HM_ERR_UNSUPPORTED;
    }
    HM_TRY(allow_lds((const void*)k_tproj_count_tiled<MT>, sizeof(MT) * (size_t)Tp));
    HM_TRY(allow_lds((const void*)k_tproj_fill_tiled<MT>, sizeof(MT) * (size_t)Tp));
    k_tproj_count_tiled<MT><<<E, kBlock, sizeof(MT) * (size_t)Tp, st>>>(
        Tp, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->voff.as<uint32_t>(),
        g->vpos.as<int32_t>(), g->lo.as<int32_t>(), g->hi.as<int32_t>(),
        g->fdeg.as<unsigned long long>(), g->bdeg.as<unsigned long long>());
  } else {
    HM_TRY(allow_lds((const void*)k_tproj_count<MT>, lds_proj));
    HM_TRY(allow_lds((const void*)k_tproj_fill<MT>, lds_proj));
    k_tproj_count<MT><<<E, kBlock, lds_proj, st>>>(
        E, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->voff.as<uint32_t>(),
        g->vpos.as<int32_t>(), g->lo.as<int32_t>(), g->hi.as<int32_t>(),
        g->fdeg.as<unsigned long long>(), g->bdeg.as<unsigned long long>());
  }
  HM_HIP(hipGetLastError());
  size_t tb1 = 0, tb2 = 0;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, g->fdeg.as<unsigned long long>(),
                                          g->foff.as<unsigned long long>(), E + 1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(
      nullptr, tb2, g->work.as<unsigned long long>(), g->work2.as<unsigned long long>(),
      g->ids.as<int32_t>(), g->order.as<int32_t>(), E, 0, 64, st));
  HM_TRY(g->tmp.ensure(std::max(tb1, tb2)));
  // fdeg[E] and bdeg[E] are the scan's trailing zero
  k_tzero<<<1, 1, 0, st>>>(g->fdeg.as<unsigned long long>() + E, 1);
  k_tzero<<<1, 1, 0, st>>>(g->bdeg.as<unsigned long long>() + E, 1);
  size_t tb = g->tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(g->tmp.p, tb, g->fdeg.as<unsigned long long>(),
                                          g->foff.as<unsigned long long>(), E + 1, st));
  tb = g->tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(g->tmp.p, tb, g->bdeg.as<unsigned long long>(),
                                          g->boff.as<unsigned long long>(), E + 1, st));
  unsigned long long tot[2] = {0, 0};
  HM_HIP(hipMemcpyAsync(&tot[0], g->foff.as<unsigned long long>() + E, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(&tot[1], g->boff.as<unsigned long long>() + E, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_TRY(g->frows.ensure(sizeof(uint32_t) * std::max<unsigned long long>(tot[0], 1)));
  HM_TRY(g->brows.ensure(sizeof(uint32_t) * std::max<unsigned long long>(tot[1], 1)));
  if (tiled) {
    const size_t nb = (size_t)std::max<unsigned long long>(tot[1], 1);
    HM_TRY(g->bmask.ensure(sizeof(MT) * nb));
    HM_TRY(g->fcur.ensure(sizeof(uint32_t) * nb));
    k_tproj_fill_tiled<MT><<<E, kBlock, sizeof(MT) * (size_t)Tp, st>>>(
        Tp, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->voff.as<uint32_t>(),
        g->vpos.as<int32_t>(), g->lo.as<int32_t>(), g->hi.as<int32_t>(),
        g->foff.as<unsigned long long>(), g->boff.as<unsigned long long>(),
        g->frows.as<uint32_t>(), g->brows.as<uint32_t>(), g->bmask.as<MT>());
  } else {
    k_tproj_fill<MT><<<E, kBlock, lds_proj, st>>>(
        E, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->voff.as<uint32_t>(),
        g->vpos.as<int32_t>(), g->lo.as<int32_t>(), g->hi.as<int32_t>(),
        g->foff.as<unsigned long long>(), g->boff.as<unsigned long long>(),
        g->frows.as<uint32_t>(), g->brows.as<uint32_t>());
  }
  HM_HIP(hipGetLastError());
  // statistics; the tiled count launches anchors most work first
  k_tstats<<<64, kBlock, 0, st>>>(E, g->lo.as<int32_t>(), g->fdeg.as<unsigned long long>(),
                                  g->bdeg.as<unsigned long long>(), g->work.as<unsigned long long>(),
                                  acc);
  if (tiled) {
    k_tiota<<<eb, kBlock, 0, st>>>(g->ids.as<int32_t>(), E);
    tb = g->tmp.bytes;
    HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(
        g->tmp.p, tb, g->work.as<unsigned long long>(), g->work2.as<unsigned long long>(),
        g->ids.as<int32_t>(), g->order.as<int32_t>(), E, 0, 64, st));
  }
  k_torient<<<(unsigned)((E + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, st>>>(
      E, g->foff.as<unsigned long long>(), g->fdeg.as<unsigned long long>(),
      g->bdeg.as<unsigned long long>(), g->frows.as<uint32_t>(), acc);
  HM_HIP(hipEventRecord(g->ev[1], st));
  if (tiled) {
    HM_TRY(allow_lds((const void*)k_tcount_tiled<MT, kTCountWaves>, sizeof(MT) * (size_t)Tc));
    k_tcount_tiled<MT, kTCountWaves><<<E, kTCountWaves * kWave, sizeof(MT) * (size_t)Tc, st>>>(
        Tc, g->eoff.as<uint32_t>(), g->enodes.as<int32_t>(), g->esize.as<uint8_t>(),
        g->voff.as<uint32_t>(), g->vpos.as<int32_t>(), g->lo.as<int32_t>(),
        g->foff.as<unsigned long long>(), g->fdeg.as<unsigned long long>(), g->frows.as<uint32_t>(),
        g->boff.as<unsigned long long>(), g->bdeg.as<unsigned long long>(), g->brows.as<uint32_t>(),
        g->bmask.as<MT>(), g->fcur.as<uint32_t>(), g->order.as<int32_t>(), acc);
    HM_HIP(hipGetLastError());
  } else {
    HM_TRY(launch_tcount<MT>(g, Wmax, lds_cnt, acc));
  }
  g->last_tile = tiled ? Tc : 0;
  HM_HIP(hipEventRecord(g->ev[2], st));
  {  // side (e) pairs of the triples open at a middle edge (both kernels take the triangles back)
    k_twedge<kWedgeWaves><<<(unsigned)((E + kWedgeWaves - 1) / kWedgeWaves), kWedgeWaves * kWave, 0, st>>>(
        E, g->lo.as<int32_t>(), g->esize.as<uint8_t>(), g->foff.as<unsigned long long>(),
        g->fdeg.as<unsigned long long>(), g->frows.as<uint32_t>(), g->boff.as<unsigned long long>(),
        g->bdeg.as<unsigned long long>(), g->brows.as<uint32_t>(), acc);
    HM_HIP(hipGetLastError());
  }
  HM_HIP(hipEventRecord(g->ev[3], st));
  HM_HIP(hipMemcpyAsync(h_acc, acc, sizeof(unsigned long long) * kTAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_HIP(hipEventElapsedTime(&g->ms_project, g->ev[0], g->ev[1]));
  HM_HIP(hipEventElapsedTime(&g->ms_count, g->ev[1], g->ev[2]));
  HM_HIP(hipEventElapsedTime(&g->ms_wedge, g->ev[2], g->ev[3]));
  return HM_OK;
}

}  // namespace hm

extern "C" {

hm_status hm_tgraph_create(const int64_t* offsets, const int32_t* nodes, const int64_t* times,
                           int32_t num_edges, hm_tgraph** out) {
  hm::clear_error();
  if (!out) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_create: null out");
  *out = nullptr;
  if (num_edges > 0 && !times) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_create: null times");
  if ((int64_t)num_edges >= hm::kMaxIds)
    return hm::fail(HM_ERR_UNSUPPORTED, "hm_tgraph_create: more than 2^22 - 1 temporal hyperedges");
  std::vector<int64_t> off;
  std::vector<int32_t> flat;
  int32_t vmax = -1;
  hm_status st = hm::prepare_host_rows(offsets, nodes, num_edges, &off, &flat, &vmax);
  if (st != HM_OK) return st;
  const int32_t E = num_edges;
  for (int32_t e = 0; e < E; ++e)
    if (times[e] < -((int64_t)1 << 62) || times[e] > ((int64_t)1 << 62))
      return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_create: timestamps must lie within +-2^62");
  hm_tgraph* g = new (std::nothrow) hm_tgraph();
  if (!g) return hm::fail(HM_ERR_OOM, "hm_tgraph_create: host allocation failed");
  auto cleanup = [&](hm_status s) { hm_tgraph_destroy(g); return s; };
  // time order, ties by input index
  std::vector<int32_t> order((size_t)E);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return times[a] < times[b]; });
  g->order_to_input = order;
  g->E = E;
  g->V = vmax + 1;
  std::vector<uint32_t> eoff((size_t)E + 1, 0), voff((size_t)g->V + 1, 0);
  std::vector<int32_t> en;
  en.reserve(flat.size());
  std::vector<uint8_t> esz((size_t)std::max(E, 1), 0);
  std::vector<long long> tpos((size_t)std::max(E, 1), 0);
  for (int32_t p = 0; p < E; ++p) {
    const int32_t e = order[(size_t)p];
    const int64_t s = off[(size_t)e], t = off[(size_t)e + 1];
    if (t - s > HM_MAX_EDGE_SIZE) {
      delete g;
      return hm::fail(HM_ERR_UNSUPPORTED, "hm_tgraph_create: hyperedge larger than 64 nodes");
    }
    for (int64_t k = s; k < t; ++k) en.push_back(flat[(size_t)k]);
    eoff[(size_t)p + 1] = (uint32_t)en.size();
    esz[(size_t)p] = (uint8_t)(t - s);
    g->max_size = std::max<int32_t>(g->max_size, (int32_t)(t - s));
    tpos[(size_t)p] = times[e];
  }
  g->I = (int64_t)en.size();
  if (g->I >= ((int64_t)1 << 32)) return cleanup(hm::fail(HM_ERR_OVER_BUDGET, "more than 2^32 incidences"));
  for (int32_t v : en) voff[(size_t)v + 1]++;
  for (int32_t v = 0; v < g->V; ++v) voff[(size_t)v + 1] += voff[(size_t)v];
  std::vector<int32_t> vp((size_t)std::max<int64_t>(g->I, 1));
  {
    std::vector<uint32_t> fill(voff.begin(), voff.end() - 1);
    for (int32_t p = 0; p < E; ++p)
      for (uint32_t k = eoff[(size_t)p]; k < eoff[(size_t)p + 1]; ++k) vp[fill[(size_t)en[k]]++] = p;
  }
  g->device = hm::current_device();
  if (hipSetDevice(g->device) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(hm::fail(HM_ERR_DEVICE, "hipStreamCreate failed"));
  for (auto& e : g->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hipEventCreate failed"));
  const size_t I1 = (size_t)std::max<int64_t>(g->I, 1);
  if ((st = g->eoff.ensure(sizeof(uint32_t) * ((size_t)E + 1))) != HM_OK ||
      (st = g->enodes.ensure(sizeof(int32_t) * I1)) != HM_OK ||
      (st = g->esize.ensure((size_t)std::max(E, 1))) != HM_OK ||
      (st = g->voff.ensure(sizeof(uint32_t) * ((size_t)g->V + 1))) != HM_OK ||
      (st = g->vpos.ensure(sizeof(int32_t) * I1)) != HM_OK ||
      (st = g->times.ensure(sizeof(long long) * (size_t)std::max(E, 1))) != HM_OK)
    return cleanup(st);
  if (hipMemcpyAsync(g->eoff.p, eoff.data(), sizeof(uint32_t) * ((size_t)E + 1), hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      (g->I > 0 && hipMemcpyAsync(g->enodes.p, en.data(), sizeof(int32_t) * (size_t)g->I, hipMemcpyHostToDevice, g->stream) != hipSuccess) ||
      (g->I > 0 && hipMemcpyAsync(g->vpos.p, vp.data(), sizeof(int32_t) * (size_t)g->I, hipMemcpyHostToDevice, g->stream) != hipSuccess) ||
      hipMemcpyAsync(g->esize.p, esz.data(), (size_t)std::max(E, 1), hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      hipMemcpyAsync(g->voff.p, voff.data(), sizeof(uint32_t) * ((size_t)g->V + 1), hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      hipMemcpyAsync(g->times.p, tpos.data(), sizeof(long long) * (size_t)std::max(E, 1), hipMemcpyHostToDevice, g->stream) != hipSuccess ||
      hipStreamSynchronize(g->stream) != hipSuccess)
    return cleanup(hm::fail(HM_ERR_DEVICE, "hm_tgraph_create: upload failed"));
  *out = g;
  return HM_OK;
}

hm_status hm_tgraph_destroy(hm_tgraph* g) {
  if (!g) return HM_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (hm::DevBuf* b : {&g->eoff, &g->enodes, &g->esize, &g->voff, &g->vpos, &g->times, &g->lo,
                        &g->hi, &g->fdeg, &g->bdeg, &g->foff, &g->boff, &g->frows, &g->brows,
                        &g->work, &g->work2, &g->ids, &g->order, &g->acc, &g->tmp, &g->bmask,
                        &g->fcur})
    b->release();
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return HM_OK;
}

hm_status hm_count_temporal(hm_tgraph* g, int64_t delta, uint64_t out[HM_TEMPORAL_BINS]) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_count_temporal: null argument");
  if (delta < 0 || delta > ((int64_t)1 << 62))
    return hm::fail(HM_ERR_BAD_INPUT, "hm_count_temporal: delta must lie in [0, 2^62]");
  unsigned long long acc[hm::kTAccCount] = {0};
  g->ms_project = g->ms_count = g->ms_wedge = 0.f;
  if (g->E >= 3) {
    if (hipSetDevice(g->device) != hipSuccess) return hm::fail(HM_ERR_DEVICE, "hipSetDevice failed");
    const hm_status st = g->max_size <= 32 ? hm::count_temporal_impl<uint32_t>(g, delta, acc)
                                           : hm::count_temporal_impl<hm::mask64_t>(g, delta, acc);
    if (st != HM_OK) return st;
  }
  for (int k = 0; k < HM_TEMPORAL_BINS; ++k) out[k] = acc[hm::kTAccHist + k];
  hm_temporal_stats& s = g->stats;
  s.num_edges = g->E;
  s.num_incidences = g->I;
  s.proj_pairs = (int64_t)acc[hm::kTAccPairs];
  s.max_window = (int64_t)acc[hm::kTAccMaxWin];
  s.neighbour_pairs = (int64_t)acc[hm::kTAccNbrPairs];
  s.oriented_row_sum = (int64_t)acc[hm::kTAccOrient];
  s.num_nodes = g->V;
  int64_t tot = 0;
  for (int k = 0; k < HM_TEMPORAL_BINS; ++k) tot += (int64_t)out[k];
  s.instances = tot;
  g->have_stats = true;
  return HM_OK;
}

hm_status hm_tgraph_set_tile(hm_tgraph* g, int32_t tile_positions) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_set_tile: null handle");
  if (tile_positions < 0) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_set_tile: negative tile");
  g->tile = tile_positions;
  return HM_OK;
}

hm_status hm_tgraph_get_stats(hm_tgraph* g, hm_temporal_stats* out) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_get_stats: null argument");
  if (!g->have_stats) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_get_stats: no count has run");
  *out = g->stats;
  return HM_OK;
}

hm_status hm_tgraph_phase_ms(hm_tgraph* g, const char* phase, double* ms) {
  hm::clear_error();
  if (!g || !phase || !ms) return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_phase_ms: null argument");
  if (std::strcmp(phase, "project") == 0) *ms = g->ms_project;
  else if (std::strcmp(phase, "count") == 0) *ms = g->ms_count;
  else if (std::strcmp(phase, "wedge") == 0) *ms = g->ms_wedge;
  else return hm::fail(HM_ERR_BAD_INPUT, "hm_tgraph_phase_ms: unknown phase");
  return HM_OK;
}

}  // extern "C"
