// Hub-aware exact static count on MI355X (gfx950): the same 30 MoCHy bins as hm_count_static
// (MoCHy-master/main_exact_par.cpp:74-173 counts every connected triple of distinct hyperedges
// once), without enumerating the cliques that frequent nodes put into the projection.
//
// The h most frequent nodes are hubs; every other node is light. Hyperedge e has the hub pattern
// A(e) (an h-bit mask) and the light part L(e). The Venn-emptiness id of a triple is the OR of
// the id its hub parts give (hub_venn over the masks) and the id its light parts give, so:
//   P0   = hub_venn | {x-only : L(x) != ∅}                (as if the light parts were disjoint)
//   P1xy = hub_venn | {xy} | {x-only : L(x) ⊄ L(y)} | {y-only : L(y) ⊄ L(x)} | {z-only : L(z) != ∅}
//          (as if only x and y shared light nodes)
//   Tx   = hub_venn | {xy, xz} | {x-only : |L(x)| > C_xy + C_xz} | {y-only : |L(y)| > C_xy}
//          | {z-only : |L(z)| > C_xz}                   (as if y and z shared no light node)
// with C_xy = |L(x) ∩ L(y)|. With F(id) the unit vector of id's bin (0 for a triple that is not
// connected), the count is the sum of four exact parts:
//   (A) classes   sum over all triples of F(P0): the triple's hyperedges enter only through their
//                 class (A(e), [L(e) != ∅]), so classes with multiplicities are enumerated;
//   (B) pairs     sum over hyperedge pairs {a, b} sharing a light node ("light pairs") and every
//                 third hyperedge c of F(P1ab) - F(P0): the pair enters through its type
//                 (A(a), [L(a) ⊄ L(b)], A(b), [L(b) ⊄ L(a)]), c through its class;
//   (C) wedges    per centre a, every pair {b, c} of its light neighbours:
//                 F(Ta) - F(P1ab) - F(P1ac) + F(P0);
//   (D) triangles every triple whose three pairs share light nodes, once:
//                 F(true) - F(P0) + F(P1ab) + F(P1bc) + F(P1ca) - F(Ta) - F(Tb) - F(Tc).
// A triple with no light pair is counted by (A) alone (P0 is its id); with one light pair {a, b}
// by (A) + (B) (P1ab is its id); with a light wedge at a by (A) + (B) + (C) (Ta is its id); with
// three light pairs by all four, which sum to its true id. Hub cliques therefore cost nothing:
// the light parts carry the enumeration (light pairs, light wedges, light triangles), the hubs
// only the class and pair-type combinatorics.
//
// Kernels:
//   k_hub_class    per hyperedge: hub mask, light size, class key
//   k_hub_classes  (A): one workgroup per class i, all class pairs j <= k at or after it
//   k_hub_centre   per hyperedge a: light mask table (a's light nodes in every x), light
//                  neighbours compacted into LDS, one pair-type key per light pair a < x written
//                  to HBM for (B), and all pairs of light neighbours for (C)
//   k_hub_types    (B): one workgroup per distinct pair type, over all classes
//   k_hub_fill     upper rows of the light projection (ascending light degree)
//   k_hub_tri      (D): oriented light triangles, as k_tri does for the full projection
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_static_dense.hpp"
#include "hm_tables.hpp"

namespace hm {

static __constant__ int8_t c_hub_bin[128] = HM_STATIC_BIN_TABLE;

constexpr int kHubIds = 128;     // slots by Venn id
constexpr int kHubCopies = 16;   // LDS copies of the slot counters (lane % 16)

__device__ inline unsigned nzm(uint32_t x) { return x != 0u ? 1u : 0u; }

// Venn-emptiness id of the hub parts of (x, y, z) (bit layout of venn_id).
__device__ inline unsigned hub_venn(uint32_t A, uint32_t B, uint32_t C) {
  return nzm(A & ~B & ~C) | nzm(B & ~A & ~C) << 1 | nzm(C & ~A & ~B) << 2 | nzm(A & B & ~C) << 3 |
         nzm(B & C & ~A) << 4 | nzm(C & A & ~B) << 5 | nzm(A & B & C) << 6;
}

__device__ inline unsigned long long binom2(unsigned long long n) {
  return (n & 1) ? n * ((n - 1) >> 1) : (n >> 1) * (n - 1);
}
// C(n, 3) mod 2^64: one of n, n - 1, n - 2 is divisible by 3 and one by 2; divide those first.
__device__ inline unsigned long long binom3(unsigned long long n) {
  if (n < 3) return 0;
  unsigned long long f[3] = {n, n - 1, n - 2};
  for (int i = 0; i < 3; ++i)
    if (f[i] % 3 == 0) { f[i] /= 3; break; }
  for (int i = 0; i < 3; ++i)
    if ((f[i] & 1) == 0) { f[i] >>= 1; break; }
  return f[0] * f[1] * f[2];
}

// Folds the workgroup's signed per-id slot counters into the 30 bins, times `mul`.
template <typename CT>
__device__ inline void fold_ids(const CT* hs, unsigned long long mul, unsigned long long* acc) {
  __shared__ unsigned long long hb[32];
  if (threadIdx.x < 32) hb[threadIdx.x] = 0ull;
  __syncthreads();
  for (int id = threadIdx.x; id < kHubIds; id += blockDim.x) {
    const int bin = c_hub_bin[id];
    if (bin < 0) continue;
    long long s = 0;
    for (int c = 0; c < kHubCopies; ++c) s += (long long)hs[id * kHubCopies + c];
    if (s) atomicAdd(&hb[bin], (unsigned long long)s * mul);
  }
  __syncthreads();
  if (threadIdx.x < kStaticBins && hb[threadIdx.x]) atomicAdd(&acc[kAccHist + threadIdx.x], hb[threadIdx.x]);
}

// Per hyperedge: hub mask, light size and class key (hub mask << 1 | [light part non-empty]).
__global__ void k_hub_class(int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
                            const int8_t* __restrict__ hub, uint32_t* __restrict__ hmask,
                            uint8_t* __restrict__ lsz, unsigned long long* __restrict__ ckey) {
  const int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= E) return;
  uint32_t m = 0;
  int l = 0;
  for (uint32_t i = eoff[e]; i < eoff[e + 1]; ++i) {
    const int hv = hub[enodes[i]];
    if (hv >= 0) m |= 1u << hv;
    else ++l;
  }
  hmask[e] = m;
  lsz[e] = (uint8_t)l;
  ckey[e] = (unsigned long long)m << 1 | (l > 0 ? 1ull : 0ull);
}

// (A): classes i <= j <= k (sorted keys, counts n); a multiset of classes stands for
// n_i n_j n_k, C(n_i, 2) n_k, n_i C(n_j, 2) or C(n_i, 3) triples, each with id P0. Classes
// before k0 have no hub (their triples are not connected through hubs) and are skipped.
__global__ __launch_bounds__(kBlock) void k_hub_classes(int32_t K, int32_t k0,
                                                        const unsigned long long* __restrict__ ukey,
                                                        const uint32_t* __restrict__ ucnt,
                                                        unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long hs[kHubIds * kHubCopies];
  const int32_t i = k0 + (int32_t)blockIdx.x;
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0ull;
  __syncthreads();
  const int lane = lane_id();
  const unsigned long long ki = ukey[i];
  const uint32_t Ai = (uint32_t)(ki >> 1);
  const unsigned li = (unsigned)(ki & 1);
  const unsigned long long ni = ucnt[i];
  for (int32_t j = i + wave_id(); j < K; j += kWavesPerBlock) {
    const unsigned long long kj = ukey[j];
    const uint32_t Aj = (uint32_t)(kj >> 1);
    const unsigned lj = (unsigned)(kj & 1);
    const unsigned long long nj = ucnt[j];
    const unsigned long long wij = (i < j) ? ni * nj : binom2(ni);      // pairs i, j (k > j)
    const unsigned long long wjj = (i < j) ? ni * binom2(nj) : binom3(ni);  // k == j
    const uint32_t Aij = Ai & Aj, Ai_j = Ai & ~Aj, Aj_i = Aj & ~Ai;
    const unsigned lo = li | lj << 1;
    for (int32_t k = j + lane; k < K; k += kWave) {
      const unsigned long long kk = ukey[k];
      const uint32_t Ak = (uint32_t)(kk >> 1);
      const unsigned lk = (unsigned)(kk & 1);
      const unsigned long long w = (k == j) ? wjj : wij * (unsigned long long)ucnt[k];
      // hub_venn(Ai, Aj, Ak) with the i, j parts precomputed
      const unsigned id = lo | nzm(Ai_j & ~Ak) | nzm(Aj_i & ~Ak) << 1 | nzm(Ak & ~(Ai | Aj)) << 2 |
                          nzm(Aij & ~Ak) << 3 | nzm(Aj_i & Ak) << 4 | nzm(Ai_j & Ak) << 5 |
                          nzm(Aij & Ak) << 6 | lk << 2;
      if (w && c_hub_bin[id] >= 0) atomicAdd(&hs[id * kHubCopies + (lane & (kHubCopies - 1))], w);
    }
  }
  __syncthreads();
  fold_ids(hs, 1ull, acc);
}

// Pair-type key of a light pair (a, x): each side is (hub mask << 1 | [its light part has a node
// the other lacks]); the larger side in the high word, so both orders give one key.
__device__ inline unsigned long long type_key(uint32_t sa, uint32_t sx) {
  const uint32_t hi = sa > sx ? sa : sx, lo = sa > sx ? sx : sa;
  return (unsigned long long)hi << 32 | lo;
}

// Key histogram of a centre's light neighbours (open addressing over 2^HB slots: in LDS with HB =
// kHubHashBits, in global memory with HB = kHubBigBits for the centres whose keys the LDS table
// cannot hold): key = hub mask << 8 | C_ax << 1 | [|L(x)| > C_ax] (hub masks of at most 23 bits),
// the neighbours with that key and those of them with a larger id. At most 7/8 of the slots fill.
constexpr int kHubHashBits = 13;
constexpr int kHubHash = 1 << kHubHashBits;
constexpr uint32_t kHubKeysMax = kHubHash - kHubHash / 8;  // lowered by HM_HUB_LDS_KEYS (tests)
// global histograms: 2^16, 2^18 and 2^20 slots, each level for the centres the one before could not
// hold, over 1024, 256 and 64 workgroups (768 MB of histograms per level)
constexpr int kHubBigLevels = 3;
constexpr int kHubBigBits[kHubBigLevels] = {16, 18, 20};
constexpr int32_t kHubBigSlices[kHubBigLevels] = {1024, 256, 64};
constexpr uint32_t kHubBigKeysMax = (1u << 20) - (1u << 20) / 8;
constexpr uint32_t kHubEmpty = 0xFFFFFFFFu;
constexpr int32_t kHubKeyHubs = 23;

// Adds a neighbour with this key; false once more than `most` (at most 7/8 of the slots) keys
// would be held.
template <int HB>
__device__ inline bool hub_key_add(uint32_t* hk, uint32_t* hn, uint32_t* hu, uint32_t key, bool up,
                                   uint32_t* used, uint32_t most) {
  constexpr uint32_t size = 1u << HB;
  uint32_t slot = (key * 2654435761u) >> (32 - HB);
  for (uint32_t p = 0; p < size; ++p) {
    const uint32_t prev = atomicCAS(&hk[slot], kHubEmpty, key);
    if (prev == kHubEmpty && atomicAdd(used, 1u) >= most) return false;
    if (prev == kHubEmpty || prev == key) {
      atomicAdd(&hn[slot], 1u);
      if (up) atomicAdd(&hu[slot], 1u);
      return true;
    }
    slot = (slot + 1) & (size - 1);
  }
  return false;
}

// a's light nodes, ascending, into s_ln (wave 0); returns nothing, callers sync before use.
__device__ inline void light_nodes(int32_t* s_ln, int32_t a, const uint32_t* __restrict__ eoff,
                                   const int32_t* __restrict__ enodes, const int8_t* __restrict__ hub) {
  if (threadIdx.x < kWave) {
    const uint32_t s = eoff[a];
    const int sa = (int)(eoff[a + 1] - s), i = (int)threadIdx.x;
    int32_t v = 0;
    bool light = false;
    if (i < sa) {
      v = enodes[s + i];
      light = hub[v] < 0;
    }
    const unsigned long long bal = __ballot(light);
    const unsigned long long below = (i == 0) ? 0ull : (~0ull >> (64 - i));
    if (light) s_ln[__popcll(bal & below)] = v;
  }
}

// Per hyperedge a: its light neighbours x over id tiles of T (the tile's mask table built from
// the parts of the V2H rows of a's light nodes that fall in it, each entry visited once by the
// lane of a's lowest light node in x), their keys in the key histogram, the light degree dL[a];
// then (C) over all pairs of keys (n_i n_j pairs of neighbours, C(n_i, 2) within one key) into
// accC, and one pair type per key with neighbours above a, with their number, appended to
// (tkeys, tcnts) for (B) (the slot count keeps counting past cap so the host can grow them).
// keys = 0: the light degree alone (no histogram, no (C), no pair types).
// HB = kHubHashBits: centres a0 + blockIdx.x, histogram in LDS (dynamic LDS: T masks, then keys,
// counts and counts above a, kHubHash each). HB > kHubHashBits: the *nin centres in in_ids, spread
// over the workgroups, each workgroup's histogram of 2^HB slots at gh (3 words per slot; dynamic
// LDS: T masks). A centre with more than `most` keys adds its light degree alone and is appended to
// out_ids[*nout], or, without out_ids, flags *overflow; once that is flagged, the later centres
// stop early (the count is redone with fewer hubs).
template <typename MT, int NW, int HB>
__global__ __launch_bounds__(NW * kWave) void k_hub_centre(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const uint32_t* __restrict__ hmask, const uint8_t* __restrict__ lsz, uint32_t* __restrict__ dL,
    unsigned long long* __restrict__ tkeys, unsigned long long* __restrict__ tcnts,
    unsigned long long* __restrict__ ntent, unsigned long long cap, uint32_t* __restrict__ overflow,
    unsigned long long* __restrict__ accC, int32_t a0, int keys, const int32_t* __restrict__ in_ids,
    const uint32_t* __restrict__ nin, int32_t* __restrict__ out_ids, uint32_t* __restrict__ nout,
    uint32_t* __restrict__ gh, uint32_t most) {
  constexpr bool BIG = HB > kHubHashBits;
  constexpr int HS = 1 << HB;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ unsigned long long hs[kHubIds * kHubCopies];
  __shared__ int32_t s_ln[HM_MAX_EDGE_SIZE];
  __shared__ TileRanges tr;
  __shared__ uint32_t s_used, s_g;
  __shared__ uint32_t s_d[NW];
  __shared__ int s_over, s_full;
  const size_t region = (size_t)T * sizeof(MT);
  MT* M = reinterpret_cast<MT*>(smem);
  uint32_t* hk = BIG ? gh + (size_t)blockIdx.x * 3 * HS : reinterpret_cast<uint32_t*>(smem + region);
  uint32_t* hn = hk + HS;
  uint32_t* hu = hn + HS;
  const int32_t nwork = BIG ? (int32_t)*nin : (int32_t)blockIdx.x + 1;
  for (int32_t it = (int32_t)blockIdx.x; it < nwork; it += BIG ? (int32_t)gridDim.x : 1) {
    __syncthreads();  // the previous centre's reads of the shared state are done
    const int32_t a = BIG ? in_ids[it] : a0 + it;
    const int la = lsz[a];
    if (la == 0) {  // no light node: no light neighbour
      if (threadIdx.x == 0) dL[a] = 0;
      continue;
    }
    // a count that will be redone stops early (one read, broadcast through LDS, so the whole
    // workgroup leaves or stays together)
    if (threadIdx.x == 0) s_over = keys ? (int)*(volatile uint32_t*)overflow : 0;
    __syncthreads();
    if (s_over) continue;
    __syncthreads();  // s_over is reset below
    light_nodes(s_ln, a, eoff, enodes, hub);
    for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0ull;
    for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
    if (keys) {
      for (int k = threadIdx.x; k < HS; k += blockDim.x) {
        hk[k] = kHubEmpty;
        hn[k] = 0u;
        hu[k] = 0u;
      }
    }
    if (threadIdx.x == 0) {
      s_used = 0u;
      s_g = 0u;
      s_over = 0;
      s_full = 0;
    }
    // A global histogram is cleared with plain stores, filled with atomics (performed in L2) and
    // read back with plain loads: a workgroup barrier orders none of these against L2 and the CU's
    // L1, so each phase ends with a device-scope fence (stores reach L2, L1 is invalidated) before
    // the barrier that starts the next one.
    if (BIG) __threadfence();
    const uint32_t Aa = hmask[a];
    uint32_t d = 0;
    for (int32_t t0 = 0; t0 < E; t0 += T) {
      const int32_t t1 = min(E, t0 + T);
      const uint32_t total = tile_ranges(tr, la, s_ln, voff, vedges, t0, t1);
      if (total == 0) continue;
      for_tile_entries(tr, la, total, vedges, [&](int i, int32_t x) {
        if (x != a) atomicOr(&M[x - t0], (MT)1 << i);
      });
      __syncthreads();
      for_tile_entries(tr, la, total, vedges, [&](int i, int32_t x) {
        if (x == a) return;
        const MT m = M[x - t0];
        if (m && low_bit(m) == i) {
          // once the histogram is full the centre only counts (it is redone, or the count fails)
          if (keys && !*(volatile int*)&s_full) {
            const int C = popc(m);
            const uint32_t key = hmask[x] << 8 | (uint32_t)C << 1 | (lsz[x] > C ? 1u : 0u);
            if (!hub_key_add<HB>(hk, hn, hu, key, x > a, &s_used, most)) s_full = 1;
          }
          ++d;
          M[x - t0] = (MT)0;
        }
      });
    }
    d = wave_sum(d);
    if (lane_id() == 0) s_d[wave_id()] = d;
    if (BIG) __threadfence();  // the histogram's atomics are done: compaction reads come from L2
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t dd = 0;
      for (int w = 0; w < NW; ++w) dd += s_d[w];
      dL[a] = dd;
      if (s_full) {
        if (out_ids) out_ids[atomicAdd(nout, 1u)] = a;
        else atomicOr(overflow, 1u);
      }
    }
    __syncthreads();                         // s_d is reused by the compaction below
    if (s_full || !keys) continue;  // uniform: read after the barriers above

    // compact the histogram in place (stable: slot k moves to the number of keys before it, never
    // past k, and a chunk is read whole before it is written); pair types of the neighbours above a
    uint32_t* lk = hk;
    uint32_t* ln = hn;
    const int lane = lane_id(), w = wave_id();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t placed = 0;
    for (int k0 = 0; k0 < HS; k0 += blockDim.x) {
      const int k = k0 + (int)threadIdx.x;
      const uint32_t key = hk[k];
      const bool have = key != kHubEmpty;
      const uint32_t n = have ? hn[k] : 0u, up = have ? hu[k] : 0u;
      const unsigned long long bal = __ballot(have);
      if (lane == 0) s_d[w] = (uint32_t)__popcll(bal);
      __syncthreads();
      uint32_t woff = 0, tot = 0;
      for (int q = 0; q < NW; ++q) {
        woff += q < w ? s_d[q] : 0;
        tot += s_d[q];
      }
      if (have) {
        const uint32_t at = placed + woff + (uint32_t)__popcll(bal & below);
        lk[at] = key;
        ln[at] = n;
      }
      placed += tot;
      const bool emit = up > 0;
      const unsigned long long eb = __ballot(emit);
      if (eb) {
        const int leader = __ffsll((long long)eb) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(ntent, (unsigned long long)__popcll(eb));
        base = __shfl(base, leader);
        if (emit) {
          const unsigned long long pos = base + (unsigned long long)__popcll(eb & below);
          if (pos < cap) {
            const int C = (int)((key >> 1) & 127u);
            const uint32_t sa = Aa << 1 | (la > C ? 1u : 0u), sx = (key >> 8) << 1 | (key & 1u);
            tkeys[pos] = type_key(sa, sx);
            tcnts[pos] = up;
          }
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) s_g = placed;
    if (BIG) __threadfence();  // the compacted keys are in L2 before the other waves read them
    __syncthreads();
    const uint32_t G = s_g;
    unsigned long long* hc = hs + (lane & (kHubCopies - 1));
    for (uint32_t i = w; i < G; i += NW) {
      const uint32_t bi = lk[i];
      const unsigned long long ni = ln[i];
      const uint32_t Ab = bi >> 8;
      const int Cb = (int)((bi >> 1) & 127u);
      const unsigned fb = bi & 1u;
      const uint32_t Aab = Aa & Ab, Aa_b = Aa & ~Ab, Ab_a = Ab & ~Aa;
      for (uint32_t j = i + lane; j < G; j += kWave) {
        const uint32_t cj = lk[j];
        const unsigned long long wt = (j == i) ? binom2(ni) : ni * (unsigned long long)ln[j];
        if (!wt) continue;
        const uint32_t Ac = cj >> 8;
        const int Cc = (int)((cj >> 1) & 127u);
        const unsigned fc = cj & 1u;
        const unsigned h = nzm(Aa_b & ~Ac) | nzm(Ab_a & ~Ac) << 1 | nzm(Ac & ~(Aa | Ab)) << 2 |
                           nzm(Aab & ~Ac) << 3 | nzm(Ab_a & Ac) << 4 | nzm(Aa_b & Ac) << 5 |
                           nzm(Aab & Ac) << 6;
        const unsigned ta = h | 8u | 32u | (la > Cb + Cc ? 1u : 0u) | fb << 1 | fc << 2;
        const unsigned p1ab = h | 8u | (la > Cb ? 1u : 0u) | fb << 1 | 4u;
        const unsigned p1ac = h | 32u | (la > Cc ? 1u : 0u) | fc << 2 | 2u;
        const unsigned p0 = h | 7u;
        atomicAdd(&hc[ta * kHubCopies], wt);
        atomicAdd(&hc[p0 * kHubCopies], wt);
        atomicAdd(&hc[p1ab * kHubCopies], 0ull - wt);
        atomicAdd(&hc[p1ac * kHubCopies], 0ull - wt);
      }
    }
    __syncthreads();
    fold_ids(hs, 1ull, accC);
  }
}

// Workgroups of one launch over a data-sized list (types, classes): a launch's work-item count
// (grid x block) is a 32-bit quantity.
constexpr unsigned long long kHubLaunch = 1ull << 20;

// (B): one workgroup per distinct pair type t = (Aa, fa, Ab, fb) with nt light pairs; every class
// (Ac, lc) with nc hyperedges contributes nc (F(P1ab) - F(P0)); the pair's own two hyperedges
// (classes (Aa, 1) and (Ab, 1)) are not third hyperedges and are taken back once. Classes before
// k0 have no hub: such a c is connected to neither a nor b, so F(P1ab) = F(P0) = 0.
__global__ __launch_bounds__(kBlock) void k_hub_types(int32_t K, int32_t k0,
                                                      const unsigned long long* __restrict__ ukey,
                                                      const uint32_t* __restrict__ ucnt,
                                                      const unsigned long long* __restrict__ tkey,
                                                      const unsigned long long* __restrict__ tcnt,
                                                      unsigned long long* __restrict__ acc, unsigned long long t0,
                                                      const uint32_t* __restrict__ list) {
  __shared__ long long hs[kHubIds * kHubCopies];
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  __syncthreads();
  // type t0 + blockIdx.x (launches of kHubLaunch types), or the list's entry there
  const unsigned long long ti = list ? (unsigned long long)list[t0 + blockIdx.x] : t0 + blockIdx.x;
  const unsigned long long t = tkey[ti];
  const uint32_t s1 = (uint32_t)(t >> 32), s2 = (uint32_t)t;
  const uint32_t Aa = s1 >> 1, Ab = s2 >> 1;
  const unsigned fa = s1 & 1u, fb = s2 & 1u;
  const uint32_t Aab = Aa & Ab, Aa_b = Aa & ~Ab, Ab_a = Ab & ~Aa;
  const int lane = lane_id();
  long long* hc = hs + (lane & (kHubCopies - 1));
  auto add_class = [&](uint32_t Ac, unsigned lc, long long nc) {
    const unsigned h = nzm(Aa_b & ~Ac) | nzm(Ab_a & ~Ac) << 1 | nzm(Ac & ~(Aa | Ab)) << 2 |
                       nzm(Aab & ~Ac) << 3 | nzm(Ab_a & Ac) << 4 | nzm(Aa_b & Ac) << 5 |
                       nzm(Aab & Ac) << 6;
    const unsigned p1 = h | 8u | fa | fb << 1 | lc << 2;
    const unsigned p0 = h | 3u | lc << 2;
    atomicAdd((unsigned long long*)&hc[p1 * kHubCopies], (unsigned long long)nc);
    atomicAdd((unsigned long long*)&hc[p0 * kHubCopies], (unsigned long long)(-nc));
  };
  for (int32_t k = k0 + (int32_t)threadIdx.x; k < K; k += blockDim.x) {
    const unsigned long long kk = ukey[k];
    add_class((uint32_t)(kk >> 1), (unsigned)(kk & 1), (long long)ucnt[k]);
  }
  if (threadIdx.x == 0) {
    add_class(Aa, 1u, -1);
    add_class(Ab, 1u, -1);
  }
  __syncthreads();
  fold_ids(hs, (unsigned long long)tcnt[ti], acc);
}

// Subset sums of the classes over hub masks, for (B) without a pass over the classes per type:
// gtab[l << h | X] = hyperedges of the classes (A, l) with A a subset of X (X < 2^h).
__global__ void k_hub_gtab_scatter(int32_t K, int hh, const unsigned long long* __restrict__ ukey,
                                   const uint32_t* __restrict__ ucnt, uint32_t* __restrict__ gtab) {
  const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (c >= K) return;
  const unsigned long long key = ukey[c];  // hub mask << 1 | [light part non-empty]; keys distinct
  gtab[((size_t)(key & 1ull) << hh) | (size_t)(key >> 1)] = ucnt[c];
}
// One subset-sum step over bit b: every mask with bit b adds the mask without it (which this step
// does not change).
__global__ void k_hub_gtab_sos(int hh, int b, uint32_t* __restrict__ gtab) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ((size_t)2 << hh)) return;
  if ((i >> b) & 1u) gtab[i] += gtab[i ^ ((size_t)1 << b)];
}

constexpr int kHubSosMaxBits = 7;  // types whose two hub masks hold more hubs go through k_hub_types

// (B) by subset sums: thread per pair type t = (Aa, fa, Ab, fb) with nt light pairs. The classes'
// contribution depends on A(c) only through S = A(c) ∩ (Aa ∪ Ab) and [A(c) has a hub outside], so
// the hyperedges per (S, outside, l) come from gtab by inclusion-exclusion over the subsets of S
// (3^|Aa ∪ Ab| lookups); each adds nt * n (F(P1ab) - F(P0)) as k_hub_types does class by class.
// Types with more than kHubSosMaxBits hubs in Aa ∪ Ab are listed for k_hub_types instead.
__global__ __launch_bounds__(kBlock) void k_hub_types_sos(unsigned long long t0, unsigned long long nT, int hh,
                                                          const unsigned long long* __restrict__ tkey,
                                                          const unsigned long long* __restrict__ tcnt,
                                                          const uint32_t* __restrict__ gtab,
                                                          uint32_t* __restrict__ slow, uint32_t* __restrict__ nslow,
                                                          unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long hs[kHubIds * kHubCopies];
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0ull;
  __syncthreads();
  unsigned long long* hc = hs + (lane_id() & (kHubCopies - 1));
  const uint32_t full = (hh >= 32) ? 0xFFFFFFFFu : ((1u << hh) - 1u);
  const uint32_t* g0 = gtab;
  const uint32_t* g1 = gtab + ((size_t)1 << hh);
  for (unsigned long long t = t0 + (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; t < t0 + nT;
       t += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long key = tkey[t];
    const uint32_t s1 = (uint32_t)(key >> 32), s2 = (uint32_t)key;
    const uint32_t Aa = s1 >> 1, Ab = s2 >> 1;
    const unsigned fa = s1 & 1u, fb = s2 & 1u;
    const uint32_t U = Aa | Ab;
    if (__popc(U) > kHubSosMaxBits) {
      slow[atomicAdd(nslow, 1u)] = (uint32_t)t;
      continue;
    }
    const unsigned long long w = tcnt[t];
    const uint32_t Ro = Aa & ~Ab, Rb = Ab & ~Aa, Rab = Aa & Ab, NU = full & ~U;
    auto add = [&](unsigned id, unsigned long long v) { atomicAdd(&hc[id * kHubCopies], v); };
    for (uint32_t S = U;; S = (S - 1u) & U) {
      long long in0 = 0, in1 = 0, all0 = 0, all1 = 0;  // classes with A ∩ U = S: inside U / any
      const int ps = __popc(S);
      for (uint32_t T = S;; T = (T - 1u) & S) {
        const long long sg = ((ps - __popc(T)) & 1) ? -1 : 1;
        in0 += sg * (long long)g0[T];
        in1 += sg * (long long)g1[T];
        all0 += sg * (long long)g0[T | NU];
        all1 += sg * (long long)g1[T | NU];
        if (T == 0u) break;
      }
      const unsigned hin = nzm(Ro & ~S) | nzm(Rb & ~S) << 1 | nzm(Rab & ~S) << 3 | nzm(Rb & S) << 4 |
                           nzm(Ro & S) << 5 | nzm(Rab & S) << 6;
      const long long n[2][2] = {{in0, all0 - in0}, {in1, all1 - in1}};  // [l][outside]
      for (unsigned l = 0; l < 2; ++l)
        for (unsigned out = 0; out < 2; ++out) {
          const long long c = n[l][out];
          if (c == 0) continue;
          const unsigned h = hin | out << 2;
          const unsigned long long v = w * (unsigned long long)c;
          add(h | 8u | fa | fb << 1 | l << 2, v);
          add(h | 3u | l << 2, 0ull - v);
        }
      if (S == 0u) break;
    }
    // the pair's own two hyperedges (classes (Aa, 1) and (Ab, 1)) are not third hyperedges
    for (int side = 0; side < 2; ++side) {
      const uint32_t Ac = side ? Ab : Aa;
      const unsigned h = nzm(Ro & ~Ac) | nzm(Rb & ~Ac) << 1 | nzm(Ac & ~U) << 2 | nzm(Rab & ~Ac) << 3 |
                         nzm(Rb & Ac) << 4 | nzm(Ro & Ac) << 5 | nzm(Rab & Ac) << 6;
      add(h | 8u | fa | fb << 1 | 4u, 0ull - w);
      add(h | 3u | 4u, w);
    }
  }
  __syncthreads();
  fold_ids(hs, 1ull, acc);
}

// Sort keys of the light-projection order: ascending light degree, with the hyperedges flagged in
// `first` (if given) ahead of all others. Any total order counts every light triangle once at its
// lowest-ranked member; ranking changed hyperedges first makes that member a changed one for every
// light triangle that holds one, so (D) over the first centres alone enumerates exactly those.
__global__ void k_hub_rank_key(int32_t E, const uint32_t* __restrict__ dL, const uint8_t* __restrict__ first,
                               uint32_t* __restrict__ key) {
  const int32_t a = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (a >= E) return;
  key[a] = (first && first[a]) ? 0u : min(dL[a], 0xFFFFFFFEu) + 1u;
}
__global__ void k_hub_flag(int32_t n, const int32_t* __restrict__ ids, uint8_t* __restrict__ flag) {
  const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < n) flag[ids[k]] = 1;
}

// Light-projection order from the sorted keys: rank of each hyperedge, hub mask, light size and
// light degree (the row capacity; degr[E] = 0 for the scan) by rank.
__global__ void k_hub_rank(int32_t E, const int32_t* __restrict__ perm, const uint32_t* __restrict__ hmask,
                           const uint8_t* __restrict__ lsz, const uint32_t* __restrict__ dL,
                           int32_t* __restrict__ rank, uint32_t* __restrict__ hmr, uint8_t* __restrict__ lszr,
                           unsigned long long* __restrict__ degr) {
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r == E) degr[E] = 0ull;
  if (r >= E) return;
  const int32_t a = perm[r];
  rank[a] = r;
  hmr[r] = hmask[a];
  lszr[r] = lsz[a];
  degr[r] = dL[a];
}

// Upper rows of the light projection (one LDS table by rank): as proj_fill_item, from a's light
// nodes, never writing past the row's capacity roff[ra + 1] - roff[ra] (the light degree the
// centre pass counted); a row that would is reported in *bad.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_hub_fill(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const uint8_t* __restrict__ lsz, const int32_t* __restrict__ rank,
    const unsigned long long* __restrict__ roff, uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows,
    uint32_t* __restrict__ bad) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t a = blockIdx.x;
  const int32_t ra = rank[a];
  if (lsz[a] == 0) {
    if (threadIdx.x == 0) rlen[ra] = 0;
    return;
  }
  MT* M = reinterpret_cast<MT*>(smem);
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, rank, hub);
  const unsigned long long base = roff[ra];
  const uint32_t cap = (uint32_t)(roff[ra + 1] - base);
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t running = 0;
  for (int32_t r0 = ra + 1; r0 < E; r0 += blockDim.x) {
    const int32_t r = r0 + (int32_t)threadIdx.x;
    const MT m = (r < E) ? M[r] : (MT)0;
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
      if (pos < cap) rows[base + pos] = ((uint32_t)r << kCntBits) | (uint32_t)popc(m);
    }
    running += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rlen[ra] = min(running, cap);
    if (running > cap) atomicOr(bad, 1u);
  }
}

// The 8 id contributions of a light triangle (roles a, b, c in ascending rank) into the int32
// slot counters hc (copy already selected); Aab, Aa_b, Ab_a are a's and b's hub-mask parts.
__device__ inline void light_triangle(int32_t* hc, uint32_t Aab, uint32_t Aa_b, uint32_t Ab_a, uint32_t Ac,
                                      int sa, int sb, int sc, int cab, int cbc, int cac, int g) {
  const unsigned h = nzm(Aa_b & ~Ac) | nzm(Ab_a & ~Ac) << 1 | nzm(Ac & ~(Aa_b | Ab_a | Aab)) << 2 |
                     nzm(Aab & ~Ac) << 3 | nzm(Ab_a & Ac) << 4 | nzm(Aa_b & Ac) << 5 | nzm(Aab & Ac) << 6;
  const unsigned tv = h | (unsigned)venn_id(sa, sb, sc, cab, cbc, cac, g);
  const unsigned p0 = h | 7u;
  const unsigned p1ab = h | 8u | (sa > cab ? 1u : 0u) | (sb > cab ? 2u : 0u) | 4u;
  const unsigned p1bc = h | 16u | (sb > cbc ? 2u : 0u) | (sc > cbc ? 4u : 0u) | 1u;
  const unsigned p1ca = h | 32u | (sc > cac ? 4u : 0u) | (sa > cac ? 1u : 0u) | 2u;
  const unsigned ta = h | 8u | 32u | (sa > cab + cac ? 1u : 0u) | (sb > cab ? 2u : 0u) | (sc > cac ? 4u : 0u);
  const unsigned tb = h | 8u | 16u | (sb > cab + cbc ? 2u : 0u) | (sa > cab ? 1u : 0u) | (sc > cbc ? 4u : 0u);
  const unsigned tc = h | 16u | 32u | (sc > cac + cbc ? 4u : 0u) | (sa > cac ? 1u : 0u) | (sb > cbc ? 2u : 0u);
  atomicAdd(&hc[tv * kHubCopies], 1);
  atomicAdd(&hc[p0 * kHubCopies], -1);
  atomicAdd(&hc[p1ab * kHubCopies], 1);
  atomicAdd(&hc[p1bc * kHubCopies], 1);
  atomicAdd(&hc[p1ca * kHubCopies], 1);
  atomicAdd(&hc[ta * kHubCopies], -1);
  atomicAdd(&hc[tb * kHubCopies], -1);
  atomicAdd(&hc[tc * kHubCopies], -1);
}

constexpr int kHubLoads = 4;  // row loads in flight per lane in the triangle kernels

// (D): centre rank r; every light triangle r < rb < rc (rb in r's upper light row, rc in rb's)
// adds F(true) - F(P0) + F(P1ab) + F(P1bc) + F(P1ca) - F(Ta) - F(Tb) - F(Tc) (roles a, b, c =
// r, rb, rc; every member has a light node). Dynamic LDS by rank: light masks M (E), hub masks
// (E words) and light sizes (E bytes). acc[kAccLightTri] counts the light triangles.
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_hub_tri(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ perm, const uint32_t* __restrict__ hmr,
    const uint8_t* __restrict__ lszr, const unsigned long long* __restrict__ roff,
    const uint32_t* __restrict__ rlen, const uint32_t* __restrict__ rows, unsigned long long* __restrict__ acc,
    int32_t r0) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t hs[kHubIds * kHubCopies];
  const int32_t r = r0 + (int32_t)blockIdx.x;
  const uint32_t nA = rlen[r];
  if (nA < 2) return;
  MT* M = reinterpret_cast<MT*>(smem);
  uint32_t* hm = reinterpret_cast<uint32_t*>(M + E);
  uint8_t* sz = reinterpret_cast<uint8_t*>(hm + E);
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  for (int x = threadIdx.x; x < E; x += blockDim.x) {
    sz[x] = lszr[x];
    hm[x] = hmr[x];
  }
  build_masks<MT>(M, E, perm[r], eoff, enodes, voff, vedges, rank, hub);

  const int lane = lane_id(), w = wave_id();
  int32_t* hc = hs + (lane & (kHubCopies - 1));
  const int sa = sz[r];
  const uint32_t Aa = hm[r];
  const uint32_t* rowA = rows + roff[r];
  unsigned long long ntri = 0;
  for (uint32_t ib = w; ib < nA; ib += NW) {
    const uint32_t eb = rowA[ib];
    const int32_t rb = (int32_t)(eb >> kCntBits);
    const int cab = (int)(eb & kCntMask);
    const MT mab = M[rb];
    const int sb = sz[rb];
    const uint32_t Ab = hm[rb];
    const uint32_t Aab = Aa & Ab, Aa_b = Aa & ~Ab, Ab_a = Ab & ~Aa;
    const uint32_t* rowB = rows + roff[rb];
    const uint32_t nB = rlen[rb];
    auto visit = [&](uint32_t ec) {
      const int32_t rc = (int32_t)(ec >> kCntBits);
      const MT mc = M[rc];
      if (mc) {
        light_triangle(hc, Aab, Aa_b, Ab_a, hm[rc], sa, sb, sz[rc], cab, (int)(ec & kCntMask), popc(mc),
                       popc((MT)(mab & mc)));
        ++ntri;
      }
    };
    uint32_t j = lane;
    for (; j + (kHubLoads - 1) * kWave < nB; j += kHubLoads * kWave) {
      uint32_t e[kHubLoads];
#pragma unroll
      for (int u = 0; u < kHubLoads; ++u) e[u] = rowB[j + u * kWave];
#pragma unroll
      for (int u = 0; u < kHubLoads; ++u) visit(e[u]);
    }
    for (; j < nB; j += kWave) visit(rowB[j]);
  }
  ntri = wave_sum(ntri);
  if (lane == 0 && ntri) atomicAdd(&acc[kAccLightTri], ntri);
  __syncthreads();
  fold_ids(hs, 1ull, acc);
}

// Upper rows of the light projection over rank tiles of T (vrank = V2H relabelled to light ranks,
// ascending per node): entries (rank << 7 | C) grouped by tile in ascending tile order, as
// k_proj_fill_tiled writes them, from a's light nodes only.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_hub_fill_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vrank, const int8_t* __restrict__ hub,
    const uint8_t* __restrict__ lsz, const int32_t* __restrict__ rank,
    const unsigned long long* __restrict__ roff, uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows,
    uint32_t* __restrict__ bad, int count_only) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ int32_t s_ln[HM_MAX_EDGE_SIZE];
  __shared__ TileRanges tr;
  __shared__ uint32_t s_cnt;
  const int32_t a = blockIdx.x;
  const int32_t ra = rank[a];
  const int la = lsz[a];
  if (la == 0) {
    if (threadIdx.x == 0) rlen[ra] = 0;
    return;
  }
  light_nodes(s_ln, a, eoff, enodes, hub);
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  if (threadIdx.x == 0) s_cnt = 0;
  uint32_t* out = count_only ? nullptr : rows + roff[ra];
  const uint32_t cap = count_only ? 0u : (uint32_t)(roff[ra + 1] - roff[ra]);
  for (int32_t t0 = ((ra + 1) / T) * T; t0 < E; t0 += T) {
    const int32_t t1 = min(E, t0 + T);
    const uint32_t total = tile_ranges(tr, la, s_ln, voff, vrank, max(t0, ra + 1), t1);
    if (total == 0) continue;
    for_tile_entries(tr, la, total, vrank, [&](int i, int32_t x) { atomicOr(&M[x - t0], (MT)1 << i); });
    __syncthreads();
    for_tile_entries(tr, la, total, vrank, [&](int i, int32_t x) {
      const MT m = M[x - t0];
      if (m && low_bit(m) == i) {
        const uint32_t pos = atomicAdd(&s_cnt, 1u);
        if (pos < cap) out[pos] = ((uint32_t)x << kCntBits) | (uint32_t)popc(m);
        M[x - t0] = (MT)0;
      }
    });
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (count_only) {  // the upper light degree alone: sizes the rows before they are written
      rlen[ra] = s_cnt;
    } else {
      rlen[ra] = min(s_cnt, cap);
      if (s_cnt > cap) atomicOr(bad, 1u);
    }
  }
}

// (D) over rank tiles of c, as k_tri_tiled walks the full projection: per b in r's upper light
// row, state at uoff[r] - ubase + ib keeps a's light mask in b and the read position in b's row (grouped
// by tile); tiles without work are skipped. The int32 slot counters are folded and cleared per
// tile (the host bounds their count per tile).
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_hub_tri_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vrank, const int8_t* __restrict__ hub,
    const int32_t* __restrict__ perm, const uint32_t* __restrict__ hmr, const uint8_t* __restrict__ lszr,
    const unsigned long long* __restrict__ roff, const uint32_t* __restrict__ rlen,
    const uint32_t* __restrict__ rows, const unsigned long long* __restrict__ uoff,
    uint32_t* __restrict__ cur, MT* __restrict__ pm, unsigned long long* __restrict__ acc, int32_t r0,
    unsigned long long ubase) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  uint32_t* hmt = reinterpret_cast<uint32_t*>(M + T);
  uint8_t* szt = reinterpret_cast<uint8_t*>(hmt + T);
  __shared__ int32_t hs[kHubIds * kHubCopies];
  __shared__ unsigned long long hb[32];
  __shared__ int32_t s_ln[HM_MAX_EDGE_SIZE];
  __shared__ TileRanges tr;
  __shared__ uint32_t s_end;
  __shared__ int32_t s_next;
  const int32_t r = r0 + (int32_t)blockIdx.x;
  const uint32_t nA = rlen[r];
  if (nA < 2) return;
  const int32_t a = perm[r];
  const int la = lszr[r];
  light_nodes(s_ln, a, eoff, enodes, hub);
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  if (threadIdx.x < 32) hb[threadIdx.x] = 0ull;
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;

  const int lane = lane_id(), w = wave_id();
  int32_t* hc = hs + (lane & (kHubCopies - 1));
  const uint32_t Aa = hmr[r];
  const uint32_t* rowA = rows + roff[r];
  const unsigned long long st = uoff[r] - ubase;  // state of this launch's centres starts at uoff[r0]
  uint32_t prev_end = 0;
  unsigned long long ntri = 0;
  __syncthreads();
  int32_t k = (int32_t)(rowA[0] >> kCntBits) / T;
  for (;;) {
    const int32_t t0 = k * T, t1 = min(E, t0 + T);
    if (threadIdx.x == 0) {  // end of the b entries whose rank is below t1 (row grouped by tile)
      uint32_t lo = prev_end, hi = nA;
      while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if ((int32_t)(rowA[mid] >> kCntBits) < t1) lo = mid + 1;
        else hi = mid;
      }
      s_end = lo;
    }
    const uint32_t total = tile_ranges(tr, la, s_ln, voff, vrank, max(t0, r + 1), t1);
    if (threadIdx.x == 0) s_next = INT32_MAX;
    for_tile_entries(tr, la, total, vrank, [&](int i, int32_t x) {
      atomicOr(&M[x - t0], (MT)1 << i);
      szt[x - t0] = lszr[x];
      hmt[x - t0] = hmr[x];
    });
    __syncthreads();
    const uint32_t end = s_end;
    for (uint32_t base = (uint32_t)w * kWave; base < end; base += NW * kWave) {
      const uint32_t ib = base + lane;
      const bool have = ib < end;
      int32_t rb = 0;
      int cab = 0;
      uint32_t cb = 0, nB = 0;
      unsigned long long ob = 0;
      MT mab = (MT)0;
      bool work = false;
      if (have) {
        const uint32_t eb = rowA[ib];
        rb = (int32_t)(eb >> kCntBits);
        cab = (int)(eb & kCntMask);
        ob = roff[rb];
        nB = rlen[rb];
        if (ib >= prev_end) {
          mab = pair_mask<MT>(s_ln, la, eoff, enodes, perm[rb]);
          pm[st + ib] = mab;
        } else {
          mab = pm[st + ib];
          cb = cur[st + ib];
        }
        work = cb < nB && (int32_t)(rows[ob + cb] >> kCntBits) < t1;
      }
      unsigned long long wl = __ballot(work);
      while (wl) {
        const int leader = __ffsll((long long)wl) - 1;
        wl &= wl - 1;
        const int32_t rbL = __shfl(rb, leader);
        const int cabL = __shfl(cab, leader);
        const uint32_t nBL = __shfl(nB, leader);
        const unsigned long long obL = __shfl(ob, leader);
        const MT mabL = __shfl(mab, leader);
        uint32_t pos = __shfl(cb, leader);
        const int sbL = lszr[rbL];
        const uint32_t AbL = hmr[rbL];
        const uint32_t Aab = Aa & AbL, Aa_b = Aa & ~AbL, Ab_a = AbL & ~Aa;
        const uint32_t* rowB = rows + obL;
        // kHubLoads chunks of b's row per step; the row is grouped by tile, so the entries inside
        // the tile are a prefix from pos and each chunk's valid entries follow the previous chunk's
        for (;;) {
          uint32_t e[kHubLoads];
          bool v[kHubLoads];
#pragma unroll
          for (int u = 0; u < kHubLoads; ++u) {
            const uint32_t j = pos + (uint32_t)(u * kWave) + lane;
            e[u] = j < nBL ? rowB[j] : 0xFFFFFFFFu;
            v[u] = j < nBL && (int32_t)(e[u] >> kCntBits) < t1;
          }
          uint32_t nv = 0;
#pragma unroll
          for (int u = 0; u < kHubLoads; ++u) {
            if (v[u]) {
              const int32_t rc = (int32_t)(e[u] >> kCntBits);
              const MT mc = M[rc - t0];
              if (mc) {
                light_triangle(hc, Aab, Aa_b, Ab_a, hmt[rc - t0], la, sbL, szt[rc - t0], cabL,
                               (int)(e[u] & kCntMask), popc(mc), popc((MT)(mabL & mc)));
                ++ntri;
              }
            }
            nv += (uint32_t)__popcll(__ballot(v[u]));
          }
          pos += nv;
          if (nv < (uint32_t)(kHubLoads * kWave)) break;
        }
        if (lane == leader) cb = pos;
      }
      if (have) {
        cur[st + ib] = cb;
        if (cb < nB) atomicMin(&s_next, (int32_t)(rows[ob + cb] >> kCntBits) / T);
      }
    }
    if (threadIdx.x == 0 && end < nA) atomicMin(&s_next, (int32_t)(rowA[end] >> kCntBits) / T);
    __syncthreads();
    for (int id = threadIdx.x; id < kHubIds; id += blockDim.x) {
      long long n = 0;
      for (int c = 0; c < kHubCopies; ++c) {
        n += hs[id * kHubCopies + c];
        hs[id * kHubCopies + c] = 0;
      }
      const int bin = c_hub_bin[id];
      if (n && bin >= 0) atomicAdd(&hb[bin], (unsigned long long)n);
    }
    __syncthreads();
    if (threadIdx.x < kStaticBins && hb[threadIdx.x]) {
      atomicAdd(&acc[kAccHist + threadIdx.x], hb[threadIdx.x]);
      hb[threadIdx.x] = 0ull;
    }
    for_tile_entries(tr, la, total, vrank, [&](int, int32_t x) { M[x - t0] = (MT)0; });
    const int32_t kn = s_next;
    prev_end = end;
    if (kn == INT32_MAX) break;
    k = kn;
  }
  ntri = wave_sum(ntri);
  if (lane == 0 && ntri) atomicAdd(&acc[kAccLightTri], ntri);
}

// V2H entries as (node << 32 | light rank of the hyperedge), sorted to relabel the rows.
__global__ void k_hub_node_keys(int32_t V, int64_t I, const uint32_t* __restrict__ voff,
                                const int32_t* __restrict__ vedges, const int32_t* __restrict__ rank,
                                unsigned long long* __restrict__ keys) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= I) return;
  int32_t lo = 0, hi = V;  // largest v with voff[v] <= k
  while (hi - lo > 1) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if ((int64_t)voff[mid] <= k) lo = mid;
    else hi = mid;
  }
  keys[k] = ((unsigned long long)lo << 32) | (unsigned long long)(uint32_t)rank[vedges[k]];
}
__global__ void k_hub_low_words(const unsigned long long* __restrict__ keys, int32_t* __restrict__ out,
                                int64_t I) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < I) out[k] = (int32_t)(uint32_t)keys[k];
}

__global__ void k_hub_zero(unsigned long long* p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0ull;
}
__global__ void k_hub_add(unsigned long long* to, const unsigned long long* from, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) to[i] += from[i];
}
__global__ void k_hub_iota(int32_t* p, int32_t n) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = i;
}
// degr[r] = sorted light degree at rank r (degr[E] = 0), for the exclusive scan into row offsets.
__global__ void k_hub_widen(const uint32_t* __restrict__ in, unsigned long long* __restrict__ out, int32_t E) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i <= E) out[i] = i < E ? (unsigned long long)in[i] : 0ull;
}

// ---------------------------------------------------------------------------------------------
// Host side

// Static LDS beside the dynamic tables: k_hub_centre's u64 slot counters (16 KB) and small
// arrays; k_hub_tri(_tiled)'s int32 counters (8 KB).
constexpr size_t kHubCentreStatic = 17 * 1024 + 1024;
constexpr size_t kHubTriStatic = 12 * 1024;
constexpr int kHubWaves = 16;  // waves per workgroup of the LDS-heavy kernels (one workgroup per CU)

static int hub_lds_limit(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
    v = 64 * 1024;
  return v;
}

// Centre tile: ids per mask table of k_hub_centre; the region also holds the compacted histogram.
template <typename MT>
static int32_t hub_centre_tile(int32_t E) { return std::min<int32_t>(E, sizeof(MT) == 4 ? 8192 : 4096); }
template <typename MT>
static size_t hub_centre_lds(int32_t T) { return (size_t)T * sizeof(MT) + (size_t)kHubHash * 12; }

// Hub count chosen by the cost model over h in {2, 4, ..., 23}: class triples K^3 / 6, plus light
// node triples sum_{v light} C(deg v, 3) (a proxy of the light triangles, 8 slot updates each) and
// light node pairs sum_{v light} C(deg v, 2) (pair types and key histograms), from the store read
// back to the host.
hm_status hub_choose(hm_graph* g, int32_t* h_out, double* cost_out) {
  DeviceCsr& c = g->csr;
  const int32_t E = c.E, V = c.V;
  std::vector<uint32_t> eoff((size_t)E + 1), voff((size_t)V + 1);
  std::vector<int32_t> enodes((size_t)std::max<int64_t>(c.I, 1));
  HM_HIP(hipMemcpyAsync(eoff.data(), c.eoff.p, sizeof(uint32_t) * (E + 1), hipMemcpyDeviceToHost, g->stream));
  HM_HIP(hipMemcpyAsync(voff.data(), c.voff.p, sizeof(uint32_t) * (V + 1), hipMemcpyDeviceToHost, g->stream));
  if (c.I > 0)
    HM_HIP(hipMemcpyAsync(enodes.data(), c.enodes.p, sizeof(int32_t) * c.I, hipMemcpyDeviceToHost, g->stream));
  HM_HIP(hipStreamSynchronize(g->stream));
  std::vector<int32_t> order((size_t)V);
  std::iota(order.begin(), order.end(), 0);
  const int32_t top = std::min<int32_t>(kHubKeyHubs, V);
  auto deg = [&](int32_t v) { return (double)(voff[v + 1] - voff[v]); };
  std::partial_sort(order.begin(), order.begin() + top, order.end(), [&](int32_t x, int32_t y) {
    const uint32_t dx = voff[x + 1] - voff[x], dy = voff[y + 1] - voff[y];
    return dx != dy ? dx > dy : x < y;
  });
  std::vector<int8_t> idx((size_t)std::max(V, 1), (int8_t)-1);
  for (int32_t i = 0; i < top; ++i) idx[order[i]] = (int8_t)i;
  double lp_all = 0, lt_all = 0;
  for (int32_t v = 0; v < V; ++v) {
    const double d = deg(v);
    lp_all += d * (d - 1) / 2;
    lt_all += d * (d - 1) * (d - 2) / 6;
  }
  int32_t best_h = 1;
  double best = -1;
  std::vector<unsigned long long> keys((size_t)E);
  for (int32_t h : {2, 4, 6, 8, 10, 12, 14, 16, 20, 23}) {
    if (h > top) break;
    double lp = lp_all, lt = lt_all;
    for (int32_t i = 0; i < h; ++i) {
      const double d = deg(order[i]);
      lp -= d * (d - 1) / 2;
      lt -= d * (d - 1) * (d - 2) / 6;
    }
    for (int32_t e = 0; e < E; ++e) {
      uint32_t m = 0;
      bool light = false;
      for (uint32_t i = eoff[e]; i < eoff[e + 1]; ++i) {
        const int32_t v = enodes[i];
        if (idx[v] >= 0 && idx[v] < h) m |= 1u << idx[v];
        else light = true;
      }
      keys[e] = (unsigned long long)m << 1 | (light ? 1ull : 0ull);
    }
    std::sort(keys.begin(), keys.end());
    const double K = (double)(std::unique(keys.begin(), keys.end()) - keys.begin());
    const double cost = K * K * K / 6 + 8 * lt + 4 * lp;
    if (best < 0 || cost < best) {
      best = cost;
      best_h = h;
    }
  }
  *h_out = best_h;
  if (cost_out) *cost_out = best;
  return HM_OK;
}

// Pair-type entries held at once (the centre pass runs in ranges of centres beyond it), and light
// rows beyond which the tiled path sizes them by the upper light degrees (one more fill pass)
// instead of the light degrees.
constexpr unsigned long long kHubTypeCap = 1ull << 28;
constexpr int32_t kHubCentreSpan = 1 << 19;  // centres of one range at most (each range reports its end)
constexpr unsigned long long kHubRowsByDegree = 1ull << 32;
// Per-launch (a, b) state of the tiled light-triangle kernel: centres are launched in ranges whose
// upper rows hold at most this many entries in all (one centre's row may exceed it alone).
constexpr unsigned long long kHubStateCap = 1ull << 30;
// The three limits above, lowered by the environment variables HM_HUB_TYPE_CAP, HM_HUB_ROWS_BY_DEGREE
// and HM_HUB_STATE_CAP (entries; tests use them to run the split paths on small graphs).
static unsigned long long hub_limit(const char* name, unsigned long long dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  char* end = nullptr;
  const unsigned long long x = std::strtoull(v, &end, 10);
  return (end && *end == '\0') ? std::min(x, dflt) : dflt;
}

// (keys, counts)[0, n) sorted by key into (k2, c2), then reduced by key into (kout, cout); the
// number of distinct keys goes to *nout (device). kout / cout may alias keys / counts.
static hm_status sort_reduce_types(CountScratch& s, unsigned long long* keys, unsigned long long* counts,
                                   unsigned long long* k2, unsigned long long* c2, unsigned long long* kout,
                                   unsigned long long* cout, unsigned long long n, int bits,
                                   unsigned long long* nout, hipStream_t st) {
  size_t need1 = 0, need2 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need1, keys, k2, counts, c2, (int)n, 0, bits, st));
  HM_HIP(hipcub::DeviceReduce::ReduceByKey(nullptr, need2, k2, kout, c2, cout, nout, hipcub::Sum(), (int)n, st));
  HM_TRY(s.tmp.ensure(std::max(need1, need2)));
  size_t tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, tb, keys, k2, counts, c2, (int)n, 0, bits, st));
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceReduce::ReduceByKey(s.tmp.p, tb, k2, kout, c2, cout, nout, hipcub::Sum(), (int)n, st));
  return HM_OK;
}

// b grown to at least need bytes with its first keep bytes kept.
static hm_status grow_keep(DevBuf& b, size_t need, size_t keep, hipStream_t st) {
  if (need <= b.bytes && b.p) return HM_OK;
  if (keep == 0) return b.ensure(need);
  DevBuf n;
  HM_TRY(n.ensure(std::max(need, b.bytes + b.bytes / 2)));
  hipError_t e = hipMemcpyAsync(n.p, b.p, keep, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    n.release();
    set_error("hub-aware count: copy of the pair types failed: %s", hipGetErrorString(e));
    return HM_ERR_DEVICE;
  }
  b.release();
  b.p = n.p;
  b.bytes = n.bytes;
  n.p = nullptr;
  n.bytes = 0;
  return HM_OK;
}

// The fill's capacity flag: a light row longer than the light degree the centre pass counted would
// mean the two passes disagree; the count stops rather than use the truncated rows.
static hm_status check_fill(const unsigned long long* flag, hipStream_t st) {
  unsigned long long f = 0;
  HM_HIP(hipMemcpyAsync(&f, flag, sizeof(f), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  if (f) return fail(HM_ERR_DEVICE, "hub-aware count: a light row outgrew its light degree");
  return HM_OK;
}

template <typename MT>
static hm_status count_static_hub_impl(hm_graph* g, int32_t h, const HubRun& run, unsigned long long* h_acc,
                                       unsigned long long* h_d, bool* key_overflow) {
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  HubScratch& hb = g->hub;
  const int32_t E = c.E, V = c.V;
  const int64_t I = c.I;
  hipStream_t st = g->stream;
  *key_overflow = false;
  const bool verbose = std::getenv("HM_VERBOSE") != nullptr;
  const auto t_start = std::chrono::steady_clock::now();
  const int lim = hub_lds_limit(g->device);
  // tiles: automatic, or forced by hm_graph_set_tile (tests and tuning; results do not change)
  const int32_t Tc = g->tile > 0 ? std::min<int32_t>(E, std::min<int32_t>(g->tile, hub_centre_tile<MT>(E)))
                                 : hub_centre_tile<MT>(E);
  const size_t lds_centre = hub_centre_lds<MT>(Tc);
  if (lds_centre + kHubCentreStatic > (size_t)lim)
    return fail(HM_ERR_UNSUPPORTED, "hub-aware count: the centre tables exceed the LDS limit");
  constexpr size_t kTriKey = sizeof(MT) + 5;  // light mask, hub mask, light size per key
  const bool dense = g->tile <= 0 && (size_t)E * kTriKey + kHubTriStatic <= (size_t)lim;
  const int32_t Tt = dense ? E
                           : g->tile > 0 ? g->tile
                                         : (int32_t)(((size_t)lim - kHubTriStatic - 1024) / kTriKey) & ~255;
  if (!dense && (size_t)Tt * kTriKey + kHubTriStatic > (size_t)lim)
    return fail(HM_ERR_BAD_INPUT, "hub-aware count: the forced tile exceeds the LDS limit");
  HM_TRY(allow_lds((const void*)k_hub_centre<MT, kHubWaves, kHubHashBits>, lds_centre));
  HM_TRY(allow_lds((const void*)k_hub_centre<MT, kHubWaves, kHubBigBits[0]>, (size_t)Tc * sizeof(MT)));
  HM_TRY(allow_lds((const void*)k_hub_centre<MT, kHubWaves, kHubBigBits[1]>, (size_t)Tc * sizeof(MT)));
  HM_TRY(allow_lds((const void*)k_hub_centre<MT, kHubWaves, kHubBigBits[2]>, (size_t)Tc * sizeof(MT)));
  if (dense) {
    HM_TRY(allow_lds((const void*)k_hub_fill<MT>, (size_t)E * sizeof(MT)));
    HM_TRY(allow_lds((const void*)k_hub_tri<MT, kHubWaves>, (size_t)E * kTriKey));
  } else {
    HM_TRY(allow_lds((const void*)k_hub_fill_tiled<MT>, (size_t)Tt * sizeof(MT)));
    HM_TRY(allow_lds((const void*)k_hub_tri_tiled<MT, kHubWaves>, (size_t)Tt * kTriKey));
  }
  if (dense ? false : I >= ((int64_t)1 << 31))
    return fail(HM_ERR_UNSUPPORTED, "hub-aware count: tiled path supports fewer than 2^31 incidences");

  // hubs: the h nodes of largest V2H degree (ties by smaller id); light pairs are bounded by the
  // sum over light nodes of C(deg, 2)
  std::vector<uint32_t> voff((size_t)V + 1);
  HM_HIP(hipMemcpyAsync(voff.data(), c.voff.p, sizeof(uint32_t) * (V + 1), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  std::vector<int32_t> order((size_t)V);
  std::iota(order.begin(), order.end(), 0);
  const int32_t hh = std::max<int32_t>(1, std::min<int32_t>(h, V));
  std::partial_sort(order.begin(), order.begin() + hh, order.end(), [&](int32_t x, int32_t y) {
    const uint32_t dx = voff[x + 1] - voff[x], dy = voff[y + 1] - voff[y];
    return dx != dy ? dx > dy : x < y;
  });
  std::vector<int8_t> hubidx((size_t)std::max(V, 1), (int8_t)-1);
  if (run.hub_nodes) {  // a pinned hub set (an update's two graphs share it)
    if ((int32_t)run.hub_nodes->size() != hh) return fail(HM_ERR_BAD_INPUT, "hub-aware count: hub set size");
    for (int32_t i = 0; i < hh; ++i) {
      const int32_t v = (*run.hub_nodes)[(size_t)i];
      if (v < 0 || v >= V) return fail(HM_ERR_BAD_INPUT, "hub-aware count: hub node out of range");
      hubidx[(size_t)v] = (int8_t)i;
    }
  } else {
    for (int32_t i = 0; i < hh; ++i) hubidx[order[i]] = (int8_t)i;
  }
  if (run.hubs_out) {
    run.hubs_out->clear();
    for (int32_t v = 0; v < V; ++v)
      if (hubidx[(size_t)v] >= 0) run.hubs_out->push_back(v);
    std::sort(run.hubs_out->begin(), run.hubs_out->end(),
              [&](int32_t x, int32_t y) { return hubidx[(size_t)x] < hubidx[(size_t)y]; });
  }
  unsigned long long lp_bound = 0, dmax = 0;
  for (int32_t v = 0; v < V; ++v) {
    const unsigned long long d = voff[v + 1] - voff[v];
    dmax = std::max(dmax, d);
    if (hubidx[v] < 0) lp_bound += d * (d - (d > 0)) / 2;
  }
  lp_bound = std::min<unsigned long long>(lp_bound, (unsigned long long)E * (unsigned long long)(E - 1) / 2);
  g->hub_stats = hm_hub_stats{};
  g->hub_stats.hubs = hh;

  HM_TRY(hb.hubidx.ensure((size_t)std::max(V, 1)));
  HM_TRY(hb.hmask.ensure(sizeof(uint32_t) * E));
  HM_TRY(hb.lsz.ensure((size_t)E));
  HM_TRY(hb.ckey.ensure(sizeof(unsigned long long) * E));
  HM_TRY(hb.ckey2.ensure(sizeof(unsigned long long) * E));
  HM_TRY(hb.ukey.ensure(sizeof(unsigned long long) * E));
  HM_TRY(hb.ucnt.ensure(sizeof(uint32_t) * E));
  HM_TRY(hb.nruns.ensure(sizeof(unsigned long long) * 8));
  HM_TRY(hb.hmr.ensure(sizeof(uint32_t) * E));
  HM_TRY(hb.lszr.ensure((size_t)E));
  HM_TRY(s.deg.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.ids.ensure(sizeof(int32_t) * E));
  HM_TRY(s.perm.ensure(sizeof(int32_t) * E));
  HM_TRY(s.rank.ensure(sizeof(int32_t) * E));
  HM_TRY(s.degr.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.roff.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.rlen.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.work2.ensure(sizeof(unsigned long long) * E));
  HM_TRY(s.acc.ensure(sizeof(unsigned long long) * kAccCount));
  HM_TRY(s.acc2.ensure(sizeof(unsigned long long) * kAccCount));
  unsigned long long* acc = s.acc.as<unsigned long long>();
  unsigned long long* accC = s.acc2.as<unsigned long long>();
  HM_TRY(hb.accd.ensure(sizeof(unsigned long long) * kAccCount));
  unsigned long long* accD = hb.accd.as<unsigned long long>();
  // runs: [0] classes, [1] pair-type entries, [2] distinct pair types, [3] key-overflow flag,
  // [4] fill capacity flag, [5] / [6] centres listed for the next key-histogram level, [7] pair
  // types listed for the class-by-class kernel
  unsigned long long* runs = hb.nruns.as<unsigned long long>();
  const unsigned eb = (unsigned)((E + kBlock - 1) / kBlock);
  const int tbits = 32 + hh + 1;

  HM_HIP(hipMemcpyAsync(hb.hubidx.p, hubidx.data(), (size_t)std::max(V, 1), hipMemcpyHostToDevice, st));
  HM_HIP(hipEventRecord(g->ev[0], st));
  k_hub_zero<<<1, 64, 0, st>>>(acc, kAccCount);
  k_hub_zero<<<1, 64, 0, st>>>(accD, kAccCount);
  k_hub_zero<<<1, 64, 0, st>>>(runs, 8);  // [4]: a light row outgrew the light degree
  k_hub_class<<<eb, kBlock, 0, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), hb.hubidx.as<int8_t>(),
                                     hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(),
                                     hb.ckey.as<unsigned long long>());
  HM_HIP(hipGetLastError());

  // (A) classes: sorted unique keys with counts
  size_t tb1 = 0, tb2 = 0, tb3 = 0, tb4 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, hb.ckey.as<unsigned long long>(),
                                           hb.ckey2.as<unsigned long long>(), E, 0, hh + 1, st));
  HM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, hb.ckey2.as<unsigned long long>(),
                                               hb.ukey.as<unsigned long long>(), hb.ucnt.as<uint32_t>(),
                                               runs, E, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  HM_TRY(s.tmp.ensure(std::max(std::max(tb1, tb2), std::max(tb3, tb4))));
  size_t tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortKeys(s.tmp.p, tb, hb.ckey.as<unsigned long long>(),
                                           hb.ckey2.as<unsigned long long>(), E, 0, hh + 1, st));
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRunLengthEncode::Encode(s.tmp.p, tb, hb.ckey2.as<unsigned long long>(),
                                               hb.ukey.as<unsigned long long>(), hb.ucnt.as<uint32_t>(),
                                               runs, E, st));
  unsigned long long K64 = 0, first_keys[2] = {0, 0};
  HM_HIP(hipMemcpyAsync(&K64, runs, sizeof(K64), hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(first_keys, hb.ukey.p, sizeof(first_keys), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  const int32_t K = (int32_t)K64;
  int32_t k0 = 0;  // classes with no hub: keys 0 and 1
  while (k0 < K && k0 < 2 && (first_keys[k0] >> 1) == 0) ++k0;
  g->hub_stats.classes = K;
  if (verbose) std::fprintf(stderr, "[hm hub] %d hubs, %d classes\n", hh, K);
  HM_HIP(hipEventRecord(g->ev[1], st));

  // light neighbours in ranges of centres: (C), light degrees, and the range's pair-type entries for
  // (B), sorted, reduced by type and merged into (tkey, tcnt) (nT distinct types). A range whose
  // entries outgrow the buffer runs again, with a buffer of the reported size while that stays
  // within kHubTypeCap entries, else as a shorter range; a range's (C) goes to accR first and into
  // accC once the range has held. Without run.abc the pass computes the light degrees alone.
  k_hub_zero<<<1, 64, 0, st>>>(accC, kAccCount);
  unsigned long long nt_all = 0, nT = 0;
  if (!run.abc) {
    k_hub_zero<<<1, 64, 0, st>>>(runs + 1, 3);
    k_hub_centre<MT, kHubWaves, kHubHashBits><<<E, kHubWaves * kWave, lds_centre, st>>>(
        E, Tc, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
        hb.hubidx.as<int8_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(), s.deg.as<uint32_t>(), nullptr,
        nullptr, runs + 1, 0ull, reinterpret_cast<uint32_t*>(runs + 3), accC, 0, 0, nullptr, nullptr, nullptr,
        nullptr, nullptr, 0u);
    HM_HIP(hipGetLastError());
  } else {
    HM_TRY(hb.accr.ensure(sizeof(unsigned long long) * kAccCount));
    HM_TRY(hb.big.ensure(sizeof(int32_t) * (size_t)E));
    HM_TRY(hb.big2.ensure(sizeof(int32_t) * (size_t)E));
    const uint32_t lds_keys = (uint32_t)std::max<unsigned long long>(1, hub_limit("HM_HUB_LDS_KEYS", kHubKeysMax));
    // keys of the global histograms before the last level (lowered by HM_HUB_BIG_KEYS in tests)
    const uint32_t big_keys = (uint32_t)std::max<unsigned long long>(1, hub_limit("HM_HUB_BIG_KEYS", kHubBigKeysMax));
    unsigned long long* accR = hb.accr.as<unsigned long long>();
    const unsigned long long type_cap = std::max<unsigned long long>(1, hub_limit("HM_HUB_TYPE_CAP", kHubTypeCap));
    unsigned long long cap = std::max<unsigned long long>(
        1, std::min<unsigned long long>({lp_bound, std::max<unsigned long long>((unsigned long long)E * 1024, 1u << 20),
                                         type_cap}));
    // one range up to 2^16 centres; beyond, the first range is an eighth of the longest and each
    // later one is sized from the entries per centre of the one before, to fill ~80% of the buffer
    int32_t a0 = 0, span = E <= (1 << 16) ? E : std::max<int32_t>(1, std::min<int32_t>(E, kHubCentreSpan) / 8);
    while (a0 < E) {
      const int32_t n = std::min(span, E - a0);
      HM_TRY(hb.types.ensure(sizeof(unsigned long long) * cap));
      HM_TRY(hb.tcraw.ensure(sizeof(unsigned long long) * cap));
      cap = std::min<unsigned long long>(std::min(hb.types.bytes, hb.tcraw.bytes) / sizeof(unsigned long long),
                                         type_cap);
      k_hub_zero<<<1, 64, 0, st>>>(accR, kAccCount);
      k_hub_zero<<<1, 64, 0, st>>>(runs + 1, 3);
      k_hub_zero<<<1, 64, 0, st>>>(runs + 5, 2);
      uint32_t* const cnt_a = reinterpret_cast<uint32_t*>(runs + 5);
      uint32_t* const cnt_b = reinterpret_cast<uint32_t*>(runs + 6);
      k_hub_centre<MT, kHubWaves, kHubHashBits><<<(unsigned)n, kHubWaves * kWave, lds_centre, st>>>(
          E, Tc, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
          hb.hubidx.as<int8_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(), s.deg.as<uint32_t>(),
          hb.types.as<unsigned long long>(), hb.tcraw.as<unsigned long long>(), runs + 1, cap,
          reinterpret_cast<uint32_t*>(runs + 3), accR, a0, 1, nullptr, nullptr, hb.big.as<int32_t>(), cnt_a,
          nullptr, lds_keys);
      HM_HIP(hipGetLastError());
      unsigned long long r2[6] = {0, 0, 0, 0, 0, 0};
      HM_HIP(hipMemcpyAsync(r2, runs + 1, sizeof(r2), hipMemcpyDeviceToHost, st));
      HM_HIP(hipStreamSynchronize(st));
      const uint32_t nbig = (uint32_t)r2[4];
      // the centres a histogram could not hold, again with the next larger one in global memory
      uint32_t nlev = nbig;
      int32_t* ids_in = hb.big.as<int32_t>();
      int32_t* ids_out = hb.big2.as<int32_t>();
      uint32_t* n_in = cnt_a;
      uint32_t* n_out = cnt_b;
      for (int lv = 0; lv < kHubBigLevels && nlev > 0; ++lv) {
        const int32_t slices = std::min<int32_t>((int32_t)nlev, kHubBigSlices[lv]);
        HM_TRY(hb.bighash.ensure(sizeof(uint32_t) * 3 * ((size_t)1 << kHubBigBits[lv]) * (size_t)slices));
        const bool last = lv + 1 == kHubBigLevels;
        k_hub_zero<<<1, 64, 0, st>>>(reinterpret_cast<unsigned long long*>(n_out), 1);
        auto launch = [&](auto lvl) {
          constexpr int HBv = decltype(lvl)::value;
          k_hub_centre<MT, kHubWaves, HBv><<<(unsigned)slices, kHubWaves * kWave, (size_t)Tc * sizeof(MT), st>>>(
              E, Tc, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
              hb.hubidx.as<int8_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(), s.deg.as<uint32_t>(),
              hb.types.as<unsigned long long>(), hb.tcraw.as<unsigned long long>(), runs + 1, cap,
              reinterpret_cast<uint32_t*>(runs + 3), accR, a0, 1, ids_in, n_in, last ? nullptr : ids_out,
              last ? nullptr : n_out, hb.bighash.as<uint32_t>(),
              last ? (1u << HBv) - (1u << HBv) / 8 : std::min<uint32_t>((1u << HBv) - (1u << HBv) / 8, big_keys));
        };
        if (lv == 0) launch(std::integral_constant<int, kHubBigBits[0]>{});
        else if (lv == 1) launch(std::integral_constant<int, kHubBigBits[1]>{});
        else launch(std::integral_constant<int, kHubBigBits[2]>{});
        HM_HIP(hipGetLastError());
        uint32_t nnext = 0;
        if (!last) {
          HM_HIP(hipMemcpyAsync(&nnext, n_out, sizeof(nnext), hipMemcpyDeviceToHost, st));
          HM_HIP(hipStreamSynchronize(st));
        }
        if (verbose)
          std::fprintf(stderr, "[hm hub] centres %d..%d: %u with more keys than %s holds, in %d-slot histograms\n",
                       a0, a0 + n, nlev, lv == 0 ? "LDS" : "the last level", 1 << kHubBigBits[lv]);
        nlev = nnext;
        std::swap(ids_in, ids_out);
        std::swap(n_in, n_out);
      }
      if (nbig > 0) {
        HM_HIP(hipMemcpyAsync(r2, runs + 1, sizeof(r2), hipMemcpyDeviceToHost, st));
        HM_HIP(hipStreamSynchronize(st));
      }
      if (r2[2]) {
        *key_overflow = true;
        set_error("hub-aware count: a hyperedge's light neighbours hold more than %u distinct (hub pattern, "
                  "overlap) keys with %d hubs; use fewer hubs", kHubBigKeysMax, hh);
        return HM_ERR_UNSUPPORTED;
      }
      const unsigned long long nt = r2[0];
      if (nt > cap) {  // the range again: a larger buffer, or fewer centres
        if (verbose)
          std::fprintf(stderr, "[hm hub] centres %d..%d: %llu pair-type entries outgrew %llu, again\n", a0, a0 + n,
                       nt, cap);
        if (nt <= type_cap) {
          cap = nt;
        } else {
          if (n == 1) {
            set_error("hub-aware count: one hyperedge has %llu pair-type entries", nt);
            return HM_ERR_UNSUPPORTED;
          }
          span = std::max<int32_t>(1, (int32_t)((double)n * (double)cap / (double)nt * 0.9));
        }
        continue;
      }
      k_hub_add<<<1, 64, 0, st>>>(accC, accR, kAccCount);
      g->hub_stats.big_centres += nbig;
      nt_all += nt;
      if (nt > 0) {  // sort and reduce the range's entries, append them, merge with the earlier ranges
        HM_TRY(hb.types2.ensure(sizeof(unsigned long long) * std::max(nt, nT)));
        HM_TRY(hb.tcsrt.ensure(sizeof(unsigned long long) * std::max(nt, nT)));
        HM_TRY(sort_reduce_types(s, hb.types.as<unsigned long long>(), hb.tcraw.as<unsigned long long>(),
                                 hb.types2.as<unsigned long long>(), hb.tcsrt.as<unsigned long long>(),
                                 hb.types.as<unsigned long long>(), hb.tcraw.as<unsigned long long>(), nt, tbits,
                                 runs + 2, st));
        unsigned long long nR = 0;
        HM_HIP(hipMemcpyAsync(&nR, runs + 2, sizeof(nR), hipMemcpyDeviceToHost, st));
        HM_HIP(hipStreamSynchronize(st));
        HM_TRY(grow_keep(hb.tkey, sizeof(unsigned long long) * (nT + nR), sizeof(unsigned long long) * nT, st));
        HM_TRY(grow_keep(hb.tcnt, sizeof(unsigned long long) * (nT + nR), sizeof(unsigned long long) * nT, st));
        HM_HIP(hipMemcpyAsync(hb.tkey.as<unsigned long long>() + nT, hb.types.p, sizeof(unsigned long long) * nR,
                              hipMemcpyDeviceToDevice, st));
        HM_HIP(hipMemcpyAsync(hb.tcnt.as<unsigned long long>() + nT, hb.tcraw.p, sizeof(unsigned long long) * nR,
                              hipMemcpyDeviceToDevice, st));
        if (nT > 0) {
          const unsigned long long nm = nT + nR;
          HM_TRY(hb.types2.ensure(sizeof(unsigned long long) * nm));
          HM_TRY(hb.tcsrt.ensure(sizeof(unsigned long long) * nm));
          HM_TRY(sort_reduce_types(s, hb.tkey.as<unsigned long long>(), hb.tcnt.as<unsigned long long>(),
                                   hb.types2.as<unsigned long long>(), hb.tcsrt.as<unsigned long long>(),
                                   hb.tkey.as<unsigned long long>(), hb.tcnt.as<unsigned long long>(), nm, tbits,
                                   runs + 2, st));
          HM_HIP(hipMemcpyAsync(&nR, runs + 2, sizeof(nR), hipMemcpyDeviceToHost, st));
          HM_HIP(hipStreamSynchronize(st));
          nT = nR;
        } else {
          nT = nR;
        }
      }
      if (verbose && n < E)
        std::fprintf(stderr, "[hm hub] light neighbours: centres %d..%d of %d done, %llu pair types\n", a0, a0 + n,
                     E, nT);
      a0 += n;
      span = (int32_t)std::min<double>(kHubCentreSpan,
                                       std::max(1.0, (double)n * 0.8 * (double)cap / (double)std::max(nt, 1ull)));
    }
  }
  g->hub_stats.light_pairs = -1;
  HM_HIP(hipEventRecord(g->ev[2], st));
  // (A) once the histograms are known to hold (class triples are the costly part at many hubs)
  if (run.abc) {
    for (int32_t i0 = k0; i0 < K; i0 += (int32_t)kHubLaunch) {  // class i0 + blockIdx.x
      const int32_t n = (int32_t)std::min<int64_t>(kHubLaunch, K - i0);
      k_hub_classes<<<(unsigned)n, kBlock, 0, st>>>(K, i0, hb.ukey.as<unsigned long long>(),
                                                    hb.ucnt.as<uint32_t>(), acc);
      HM_HIP(hipGetLastError());
    }
  }
  if (verbose) std::fprintf(stderr, "[hm hub] classes and light neighbours done, %llu pair-type entries\n", nt_all);

  // (B) every distinct pair type against every class
  g->hub_stats.pair_types = (int64_t)nT;
  if (run.abc && K > k0 && nT > 0) {
    // subset sums of the classes over hub masks, then the types in chunks: those with few hubs by
    // the sums, the others listed and run against every class
    HM_TRY(hb.gtab.ensure(sizeof(uint32_t) * ((size_t)2 << hh)));
    HM_HIP(hipMemsetAsync(hb.gtab.p, 0, sizeof(uint32_t) * ((size_t)2 << hh), st));
    k_hub_gtab_scatter<<<(unsigned)((K + kBlock - 1) / kBlock), kBlock, 0, st>>>(
        K, hh, hb.ukey.as<unsigned long long>(), hb.ucnt.as<uint32_t>(), hb.gtab.as<uint32_t>());
    HM_HIP(hipGetLastError());
    for (int b = 0; b < hh; ++b) {
      k_hub_gtab_sos<<<(unsigned)((((size_t)2 << hh) + kBlock - 1) / kBlock), kBlock, 0, st>>>(
          hh, b, hb.gtab.as<uint32_t>());
      HM_HIP(hipGetLastError());
    }
    const unsigned long long chunk = kHubLaunch * 4;
    HM_TRY(hb.slow.ensure(sizeof(uint32_t) * (size_t)std::min(nT, chunk)));
    uint32_t* nslow = reinterpret_cast<uint32_t*>(runs + 7);
    for (unsigned long long t0 = 0; t0 < nT; t0 += chunk) {
      const unsigned long long n = std::min(chunk, nT - t0);
      k_hub_zero<<<1, 64, 0, st>>>(runs + 7, 1);
      k_hub_types_sos<<<(unsigned)std::min<unsigned long long>((n + kBlock - 1) / kBlock, 4096), kBlock, 0, st>>>(
          t0, n, hh, hb.tkey.as<unsigned long long>(), hb.tcnt.as<unsigned long long>(), hb.gtab.as<uint32_t>(),
          hb.slow.as<uint32_t>(), nslow, acc);
      HM_HIP(hipGetLastError());
      uint32_t ns = 0;
      HM_HIP(hipMemcpyAsync(&ns, nslow, sizeof(ns), hipMemcpyDeviceToHost, st));
      HM_HIP(hipStreamSynchronize(st));
      for (unsigned long long s0 = 0; s0 < ns; s0 += kHubLaunch) {
        const unsigned long long m = std::min<unsigned long long>(kHubLaunch, ns - s0);
        k_hub_types<<<(unsigned)m, kBlock, 0, st>>>(K, k0, hb.ukey.as<unsigned long long>(), hb.ucnt.as<uint32_t>(),
                                                    hb.tkey.as<unsigned long long>(), hb.tcnt.as<unsigned long long>(),
                                                    acc, s0, hb.slow.as<uint32_t>());
        HM_HIP(hipGetLastError());
      }
    }
  }
  if (verbose) {
    HM_HIP(hipStreamSynchronize(st));
    std::fprintf(stderr, "[hm hub] class triples and %llu pair types x %d classes done\n", nT, K);
  }
  HM_HIP(hipEventRecord(g->ev[3], st));

  // (D) light triangles over the light projection, ranked by ascending light degree (the ids in
  // run.first ahead of all others)
  const uint8_t* firstflag = nullptr;
  if (run.n_first > 0) {
    HM_TRY(hb.first.ensure((size_t)E));
    HM_TRY(hb.firstids.ensure(sizeof(int32_t) * (size_t)run.n_first));
    HM_HIP(hipMemsetAsync(hb.first.p, 0, (size_t)E, st));
    HM_HIP(hipMemcpyAsync(hb.firstids.p, run.first, sizeof(int32_t) * (size_t)run.n_first, hipMemcpyHostToDevice, st));
    k_hub_flag<<<(unsigned)((run.n_first + kBlock - 1) / kBlock), kBlock, 0, st>>>(
        run.n_first, hb.firstids.as<int32_t>(), hb.first.as<uint8_t>());
    firstflag = hb.first.as<uint8_t>();
  }
  HM_TRY(hb.rkey.ensure(sizeof(uint32_t) * (size_t)E));
  k_hub_rank_key<<<eb, kBlock, 0, st>>>(E, s.deg.as<uint32_t>(), firstflag, hb.rkey.as<uint32_t>());
  k_hub_iota<<<eb, kBlock, 0, st>>>(s.ids.as<int32_t>(), E);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, tb, hb.rkey.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  k_hub_rank<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      E, s.perm.as<int32_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(), s.deg.as<uint32_t>(),
      s.rank.as<int32_t>(), hb.hmr.as<uint32_t>(), hb.lszr.as<uint8_t>(), s.degr.as<unsigned long long>());
  const int32_t nD = run.only_first ? run.n_first : E;  // centres of (D)
  const int32_t d0 = std::min<int64_t>(std::max<int64_t>(run.d_begin, 0), nD);
  const int32_t d1 = run.d_end < 0 ? nD : (int32_t)std::min<int64_t>(std::max<int64_t>(run.d_end, d0), nD);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  unsigned long long total = 0;
  HM_HIP(hipMemcpyAsync(&total, s.roff.as<unsigned long long>() + E, sizeof(total), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  g->hub_stats.light_pairs = (int64_t)(total / 2);
  if (verbose) std::fprintf(stderr, "[hm hub] pair types done; %llu light pairs\n", total / 2);
  const unsigned long long rows_by_degree = hub_limit("HM_HUB_ROWS_BY_DEGREE", kHubRowsByDegree);
  if (total > rows_by_degree) hb.release_types();  // (B) is done; the rows need the memory
  if (dense) {
    HM_TRY(s.rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total, 1)));
    k_hub_fill<MT><<<E, kBlock, (size_t)E * sizeof(MT), st>>>(
        E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
        hb.hubidx.as<int8_t>(), hb.lsz.as<uint8_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
        s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(), reinterpret_cast<uint32_t*>(runs + 4));
    HM_HIP(hipGetLastError());
    HM_TRY(check_fill(runs + 4, st));
    HM_HIP(hipEventRecord(g->ev[5], st));
    if (d1 > d0) {
      k_hub_tri<MT, kHubWaves><<<(unsigned)(d1 - d0), kHubWaves * kWave, (size_t)E * kTriKey, st>>>(
          E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
          hb.hubidx.as<int8_t>(), s.rank.as<int32_t>(), s.perm.as<int32_t>(), hb.hmr.as<uint32_t>(),
          hb.lszr.as<uint8_t>(), s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),
          accD, d0);
      HM_HIP(hipGetLastError());
    }
    if (run.d_done) *run.d_done = d1;
  } else {
    // V2H relabelled to light ranks, ascending within every node's row
    HM_TRY(s.vkeys.ensure(sizeof(unsigned long long) * std::max<int64_t>(I, 1)));
    HM_TRY(s.vkeys2.ensure(sizeof(unsigned long long) * std::max<int64_t>(I, 1)));
    HM_TRY(s.vrank.ensure(sizeof(int32_t) * std::max<int64_t>(I, 1)));
    HM_TRY(s.uoff.ensure(sizeof(unsigned long long) * (E + 1)));
    int vbits = 1;
    while (vbits < 31 && ((int64_t)1 << vbits) <= V) ++vbits;
    size_t need = 0;
    HM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, need, s.vkeys.as<unsigned long long>(),
                                             s.vkeys2.as<unsigned long long>(), (int)I, 0, 32 + vbits, st));
    HM_TRY(s.tmp.ensure(need));
    if (I > 0) {
      const unsigned ib = (unsigned)((I + kBlock - 1) / kBlock);
      k_hub_node_keys<<<ib, kBlock, 0, st>>>(V, I, c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                             s.rank.as<int32_t>(), s.vkeys.as<unsigned long long>());
      tb = s.tmp.bytes;
      HM_HIP(hipcub::DeviceRadixSort::SortKeys(s.tmp.p, tb, s.vkeys.as<unsigned long long>(),
                                               s.vkeys2.as<unsigned long long>(), (int)I, 0, 32 + vbits, st));
      k_hub_low_words<<<ib, kBlock, 0, st>>>(s.vkeys2.as<unsigned long long>(), s.vrank.as<int32_t>(), I);
    }
    if (total > rows_by_degree) {  // rows by the upper light degree: a counting pass first
      k_hub_fill_tiled<MT><<<E, kBlock, (size_t)Tt * sizeof(MT), st>>>(
          E, Tt, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), s.vrank.as<int32_t>(),
          hb.hubidx.as<int8_t>(), hb.lsz.as<uint8_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
          s.rlen.as<uint32_t>(), nullptr, reinterpret_cast<uint32_t*>(runs + 4), 1);
      HM_HIP(hipGetLastError());
      k_hub_widen<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
          s.rlen.as<uint32_t>(), s.degr.as<unsigned long long>(), E);
      tb = s.tmp.bytes;
      HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                              s.roff.as<unsigned long long>(), E + 1, st));
      HM_HIP(hipMemcpyAsync(&total, s.roff.as<unsigned long long>() + E, sizeof(total), hipMemcpyDeviceToHost, st));
      HM_HIP(hipStreamSynchronize(st));
      if (verbose) std::fprintf(stderr, "[hm hub] upper light rows: %llu entries\n", total);
    }
    HM_TRY(s.rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total, 1)));
    k_hub_fill_tiled<MT><<<E, kBlock, (size_t)Tt * sizeof(MT), st>>>(
        E, Tt, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), s.vrank.as<int32_t>(),
        hb.hubidx.as<int8_t>(), hb.lsz.as<uint8_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
        s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(), reinterpret_cast<uint32_t*>(runs + 4), 0);
    HM_HIP(hipGetLastError());
    HM_TRY(check_fill(runs + 4, st));
    // per (a, b) state at the scan of the upper-row lengths; the int32 slot counters of a tile
    // take at most d+_max (T / 16 + 2) updates per slot and copy
    std::vector<uint32_t> rl((size_t)E);
    HM_HIP(hipMemcpyAsync(rl.data(), s.rlen.p, sizeof(uint32_t) * E, hipMemcpyDeviceToHost, st));
    HM_HIP(hipStreamSynchronize(st));
    std::vector<unsigned long long> pre((size_t)E + 1, 0ull);
    unsigned long long umax = 0;
    for (int32_t r = 0; r < E; ++r) {
      pre[(size_t)r + 1] = pre[(size_t)r] + rl[(size_t)r];
      umax = std::max<unsigned long long>(umax, rl[(size_t)r]);
    }
    if (umax * ((unsigned long long)Tt / kHubCopies + 4 * kHubLoads + 1) >= (1ull << 31))
      return fail(HM_ERR_UNSUPPORTED, "hub-aware count: an upper light row is too long for the tile counters");
    k_hub_widen<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
        s.rlen.as<uint32_t>(), s.degr.as<unsigned long long>(), E);
    tb = s.tmp.bytes;
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                            s.uoff.as<unsigned long long>(), E + 1, st));
    // centres in launches whose (a, b) state stays within kHubStateCap entries (a lone centre may
    // exceed it); with HM_VERBOSE set, also at most a sixteenth of the centres per launch, each
    // reported as it ends; with a time limit, at most 1/256 of them, the limit checked after each
    const int32_t most = run.seconds > 0 ? std::max<int32_t>(1024, nD / 256)
                         : verbose       ? std::max<int32_t>(4096, nD / 16)
                                         : std::max<int32_t>(nD, 1);
    const unsigned long long state_cap = hub_limit("HM_HUB_STATE_CAP", kHubStateCap);
    std::vector<int32_t> cut{d0};
    unsigned long long smax = 0;
    while (cut.back() < d1) {
      const int32_t r0 = cut.back();
      int32_t r1 = r0 + 1;
      while (r1 < d1 && r1 - r0 < most && pre[(size_t)r1 + 1] - pre[(size_t)r0] <= state_cap) ++r1;
      smax = std::max(smax, pre[(size_t)r1] - pre[(size_t)r0]);
      cut.push_back(r1);
    }
    HM_TRY(s.cur.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(smax, 1)));
    HM_TRY(s.pm.ensure(sizeof(MT) * (size_t)std::max<unsigned long long>(smax, 1)));
    HM_HIP(hipEventRecord(g->ev[5], st));
    if (run.d_done) *run.d_done = d1;
    for (size_t q = 0; q + 1 < cut.size(); ++q) {
      const int32_t r0 = cut[q], n = cut[q + 1] - cut[q];
      k_hub_tri_tiled<MT, kHubWaves><<<(unsigned)n, kHubWaves * kWave, (size_t)Tt * kTriKey, st>>>(
          E, Tt, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), s.vrank.as<int32_t>(),
          hb.hubidx.as<int8_t>(), s.perm.as<int32_t>(), hb.hmr.as<uint32_t>(), hb.lszr.as<uint8_t>(),
          s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),
          s.uoff.as<unsigned long long>(), s.cur.as<uint32_t>(), s.pm.as<MT>(), accD, r0, pre[(size_t)r0]);
      HM_HIP(hipGetLastError());
      if (verbose || run.seconds > 0) HM_HIP(hipStreamSynchronize(st));
      if (verbose) std::fprintf(stderr, "[hm hub] light triangles: centres %d..%d of %d done\n", r0, r0 + n, nD);
      if (run.seconds > 0 && q + 2 < cut.size() &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() >= run.seconds) {
        if (run.d_done) *run.d_done = cut[q + 1];
        break;
      }
    }
  }
  HM_HIP(hipEventRecord(g->ev[4], st));  // the triangle kernel ends the phase
  unsigned long long hC[kAccCount];
  HM_HIP(hipMemcpyAsync(h_acc, acc, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(hC, accC, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipMemcpyAsync(h_d, accD, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < kStaticBins; ++k) h_acc[kAccHist + k] += hC[kAccHist + k];
  // projection statistics from the projected degrees alone (the wedge identity
  // sum(open) + 3 sum(closed) = sum_a C(d_a, 2) checks the bins against them)
  if (run.stats) {
    HM_TRY(projection_stats(g, h_acc));
    if (verbose) std::fprintf(stderr, "[hm hub] projection statistics done\n");
  }
  (void)dmax;
  g->hub_stats.light_triangles = (int64_t)h_d[kAccLightTri];
  float m01 = 0, m12 = 0, m23 = 0, m34 = 0, m54 = 0;
  HM_HIP(hipEventElapsedTime(&m54, g->ev[5], g->ev[4]));
  g->hub_stats.ms_tri_kernel = m54;
  HM_HIP(hipEventElapsedTime(&m01, g->ev[0], g->ev[1]));
  HM_HIP(hipEventElapsedTime(&m12, g->ev[1], g->ev[2]));
  HM_HIP(hipEventElapsedTime(&m23, g->ev[2], g->ev[3]));
  HM_HIP(hipEventElapsedTime(&m34, g->ev[3], g->ev[4]));
  g->hub_stats.ms_classes = m01;
  g->hub_stats.ms_wedges = m12;
  g->hub_stats.ms_types = m23;
  g->hub_stats.ms_triangles = m34;
  g->ms_project = m01 + m12 + m23;
  g->ms_triangles = m34;
  return HM_OK;
}

hm_status hub_parts(hm_graph* g, int32_t h, const HubRun& run, unsigned long long* h_abc, unsigned long long* h_d,
                    bool* key_overflow) {
  if (h < 1 || h > kHubKeyHubs) return fail(HM_ERR_BAD_INPUT, "hub count outside 1..23");
  return g->csr.max_size > 32 ? count_static_hub_impl<mask64_t>(g, h, run, h_abc, h_d, key_overflow)
                              : count_static_hub_impl<uint32_t>(g, h, run, h_abc, h_d, key_overflow);
}

hm_status count_static_hub(hm_graph* g, int32_t h, unsigned long long* h_acc, bool* key_overflow, bool stats) {
  HubRun run;
  run.hubs_out = &g->hub_nodes;
  run.stats = stats;
  unsigned long long d[kAccCount];
  HM_TRY(hub_parts(g, h, run, h_acc, d, key_overflow));
  for (int k = 0; k < kStaticBins; ++k) h_acc[kAccHist + k] += d[kAccHist + k];
  return HM_OK;
}

}  // namespace hm
