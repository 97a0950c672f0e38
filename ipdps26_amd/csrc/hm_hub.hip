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

// Per hyperedge a: light mask table M[x] (a's light nodes contained in x), its light neighbours
// compacted into LDS as (hub mask, C_ax | [|L(x)| > C_ax] << 7), the light degree dL[a], one
// pair-type key per light neighbour x > a appended to types[] (for (B); the slot count keeps
// counting past cap so the host can grow the buffer), and (C) over all pairs of light neighbours.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_hub_centre(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const uint32_t* __restrict__ hmask, const uint8_t* __restrict__ lsz, uint32_t* __restrict__ dL,
    unsigned long long* __restrict__ types, unsigned long long* __restrict__ ntypes,
    unsigned long long cap, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t hs[kHubIds * kHubCopies];
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t a = blockIdx.x;
  const int la = lsz[a];
  if (la == 0) {  // no light node: no light neighbour
    if (threadIdx.x == 0) dL[a] = 0;
    return;
  }
  MT* M = reinterpret_cast<MT*>(smem);
  uint2* list = reinterpret_cast<uint2*>(smem + (((size_t)E * sizeof(MT) + 7) & ~(size_t)7));
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, nullptr, hub);

  const uint32_t Aa = hmask[a];
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t running = 0;
  for (int32_t x0 = 0; x0 < E; x0 += blockDim.x) {
    const int32_t x = x0 + (int32_t)threadIdx.x;
    const MT m = (x < E) ? M[x] : (MT)0;
    const unsigned long long bal = __ballot(m != (MT)0);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) {
      woff += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    uint32_t sx = 0, Ax = 0;
    int C = 0;
    if (m) {
      C = popc(m);
      Ax = hmask[x];
      const uint32_t f = lsz[x] > C ? 1u : 0u;
      list[running + woff + (uint32_t)__popcll(bal & below)] = make_uint2(Ax, (uint32_t)C | f << 7);
      sx = Ax << 1 | f;
    }
    const bool emit = m && x > a;
    const unsigned long long eb = __ballot(emit);
    if (eb) {
      const int leader = __ffsll((long long)eb) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(ntypes, (unsigned long long)__popcll(eb));
      base = __shfl(base, leader);
      if (emit) {
        const unsigned long long pos = base + (unsigned long long)__popcll(eb & below);
        if (pos < cap) types[pos] = type_key(Aa << 1 | (la > C ? 1u : 0u), sx);
      }
    }
    running += tot;
    __syncthreads();
  }
  const uint32_t n = running;
  int32_t* hc = hs + (lane & (kHubCopies - 1));
  for (uint32_t i = w; i < n; i += kWavesPerBlock) {
    const uint2 bi = list[i];
    const uint32_t Ab = bi.x;
    const int Cb = (int)(bi.y & 127u);
    const unsigned fb = bi.y >> 7;
    const uint32_t Aab = Aa & Ab, Aa_b = Aa & ~Ab, Ab_a = Ab & ~Aa;
    for (uint32_t j = i + 1 + lane; j < n; j += kWave) {
      const uint2 cj = list[j];
      const uint32_t Ac = cj.x;
      const int Cc = (int)(cj.y & 127u);
      const unsigned fc = cj.y >> 7;
      const unsigned h = nzm(Aa_b & ~Ac) | nzm(Ab_a & ~Ac) << 1 | nzm(Ac & ~(Aa | Ab)) << 2 |
                         nzm(Aab & ~Ac) << 3 | nzm(Ab_a & Ac) << 4 | nzm(Aa_b & Ac) << 5 |
                         nzm(Aab & Ac) << 6;
      const unsigned ta = h | 8u | 32u | (la > Cb + Cc ? 1u : 0u) | fb << 1 | fc << 2;
      const unsigned p1ab = h | 8u | (la > Cb ? 1u : 0u) | fb << 1 | 4u;
      const unsigned p1ac = h | 32u | (la > Cc ? 1u : 0u) | fc << 2 | 2u;
      const unsigned p0 = h | 7u;
      atomicAdd(&hc[ta * kHubCopies], 1);
      atomicAdd(&hc[p0 * kHubCopies], 1);
      atomicAdd(&hc[p1ab * kHubCopies], -1);
      atomicAdd(&hc[p1ac * kHubCopies], -1);
    }
  }
  __syncthreads();
  fold_ids(hs, 1ull, acc);
  if (threadIdx.x == 0) dL[a] = n;
}

// (B): one workgroup per distinct pair type t = (Aa, fa, Ab, fb) with nt light pairs; every class
// (Ac, lc) with nc hyperedges contributes nc (F(P1ab) - F(P0)); the pair's own two hyperedges
// (classes (Aa, 1) and (Ab, 1)) are not third hyperedges and are taken back once. Classes before
// k0 have no hub: such a c is connected to neither a nor b, so F(P1ab) = F(P0) = 0.
__global__ __launch_bounds__(kBlock) void k_hub_types(int32_t K, int32_t k0,
                                                      const unsigned long long* __restrict__ ukey,
                                                      const uint32_t* __restrict__ ucnt,
                                                      const unsigned long long* __restrict__ tkey,
                                                      const uint32_t* __restrict__ tcnt,
                                                      unsigned long long* __restrict__ acc) {
  __shared__ long long hs[kHubIds * kHubCopies];
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  __syncthreads();
  const unsigned long long t = tkey[blockIdx.x];
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
  fold_ids(hs, (unsigned long long)tcnt[blockIdx.x], acc);
}

// Light-projection order: rank by ascending light degree (sorted pairs), hub mask and light size
// by rank.
__global__ void k_hub_rank(int32_t E, const int32_t* __restrict__ perm, const uint32_t* __restrict__ hmask,
                           const uint8_t* __restrict__ lsz, int32_t* __restrict__ rank,
                           uint32_t* __restrict__ hmr, uint8_t* __restrict__ lszr) {
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= E) return;
  const int32_t a = perm[r];
  rank[a] = r;
  hmr[r] = hmask[a];
  lszr[r] = lsz[a];
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_hub_fill(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const uint8_t* __restrict__ lsz, const int32_t* __restrict__ rank,
    const unsigned long long* __restrict__ roff, uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t a = blockIdx.x;
  if (lsz[a] == 0) {
    if (threadIdx.x == 0) rlen[rank[a]] = 0;
    return;
  }
  proj_fill_item<MT>(reinterpret_cast<MT*>(smem), E, a, eoff, enodes, voff, vedges, rank, roff, rlen,
                     rows, hub);
}

// (D): centre rank r; every light triangle r < rb < rc (rb in r's upper light row, rc in rb's)
// adds F(true) - F(P0) + F(P1ab) + F(P1bc) + F(P1ca) - F(Ta) - F(Tb) - F(Tc) (roles a, b, c =
// r, rb, rc; every member has a light node). M (light masks by rank) spans E ranks; the E light
// sizes by rank follow it.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_hub_tri(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges, const int8_t* __restrict__ hub,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ perm, const uint32_t* __restrict__ hmr,
    const uint8_t* __restrict__ lszr, const unsigned long long* __restrict__ roff,
    const uint32_t* __restrict__ rlen, const uint32_t* __restrict__ rows, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t hs[kHubIds * kHubCopies];
  const int32_t r = blockIdx.x;
  const uint32_t nA = rlen[r];
  if (nA < 2) return;
  MT* M = reinterpret_cast<MT*>(smem);
  uint8_t* sz = reinterpret_cast<uint8_t*>(M + E);
  for (int k = threadIdx.x; k < kHubIds * kHubCopies; k += blockDim.x) hs[k] = 0;
  for (int x = threadIdx.x; x < E; x += blockDim.x) sz[x] = lszr[x];
  build_masks<MT>(M, E, perm[r], eoff, enodes, voff, vedges, rank, hub);

  const int lane = lane_id(), w = wave_id();
  int32_t* hc = hs + (lane & (kHubCopies - 1));
  const int sa = sz[r];
  const uint32_t Aa = hmr[r];
  const uint32_t* rowA = rows + roff[r];
  for (uint32_t ib = w; ib < nA; ib += kWavesPerBlock) {
    const uint32_t eb = rowA[ib];
    const int32_t rb = (int32_t)(eb >> kCntBits);
    const int cab = (int)(eb & kCntMask);
    const MT mab = M[rb];
    const int sb = sz[rb];
    const uint32_t Ab = hmr[rb];
    const uint32_t Aab = Aa & Ab, Aa_b = Aa & ~Ab, Ab_a = Ab & ~Aa;
    const uint32_t* rowB = rows + roff[rb];
    const uint32_t nB = rlen[rb];
    for (uint32_t j = lane; j < nB; j += kWave) {
      const uint32_t ec = rowB[j];
      const int32_t rc = (int32_t)(ec >> kCntBits);
      const MT mc = M[rc];
      if (!mc) continue;
      const int cbc = (int)(ec & kCntMask);
      const int cac = popc(mc);
      const int g = popc((MT)(mab & mc));
      const int sc = sz[rc];
      const uint32_t Ac = hmr[rc];
      const unsigned h = nzm(Aa_b & ~Ac) | nzm(Ab_a & ~Ac) << 1 | nzm(Ac & ~(Aa | Ab)) << 2 |
                         nzm(Aab & ~Ac) << 3 | nzm(Ab_a & Ac) << 4 | nzm(Aa_b & Ac) << 5 |
                         nzm(Aab & Ac) << 6;
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
  }
  __syncthreads();
  fold_ids(hs, 1ull, acc);
}

__global__ void k_hub_zero(unsigned long long* p, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) p[i] = 0ull;
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

// Dynamic LDS of k_hub_centre (mask table + 8-byte neighbour list) and of k_hub_tri (mask table
// + light sizes by rank); the static slot counters (8 KB) and fold bins sit beside them.
template <typename MT>
static size_t hub_centre_lds(int32_t E) { return (((size_t)E * sizeof(MT) + 7) & ~(size_t)7) + (size_t)E * 8; }
template <typename MT>
static size_t hub_tri_lds(int32_t E) { return (size_t)E * (sizeof(MT) + 1); }
constexpr size_t kHubStaticLds = 12 * 1024;

static int hub_lds_limit(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
    v = 64 * 1024;
  return v;
}

bool hub_fits(const hm_graph* g, int device) {
  const int lim = hub_lds_limit(device);
  const int32_t E = g->csr.E;
  if (g->csr.max_size > 32)
    return hub_centre_lds<mask64_t>(E) + kHubStaticLds <= (size_t)lim &&
           hub_tri_lds<mask64_t>(E) + kHubStaticLds <= (size_t)lim;
  return hub_centre_lds<uint32_t>(E) + kHubStaticLds <= (size_t)lim &&
         hub_tri_lds<uint32_t>(E) + kHubStaticLds <= (size_t)lim;
}

// Hub count chosen by the cost model: the h in {1, 2, 4, ..., 31} minimising class triples
// K^3 / 6 + light wedges (bound: C(sum over e's light nodes of (deg - 1), 2) per hyperedge) +
// pair types (bound: min(light pairs, 4 K^2)) x K, from the store read back to the host.
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
  const int32_t top = std::min<int32_t>(kHubMax, V);
  auto deg = [&](int32_t v) { return (double)(voff[v + 1] - voff[v]); };
  std::partial_sort(order.begin(), order.begin() + top, order.end(), [&](int32_t x, int32_t y) {
    const uint32_t dx = voff[x + 1] - voff[x], dy = voff[y + 1] - voff[y];
    return dx != dy ? dx > dy : x < y;
  });
  std::vector<int8_t> idx((size_t)std::max(V, 1), (int8_t)-1);
  for (int32_t i = 0; i < top; ++i) idx[order[i]] = (int8_t)i;
  double lp_all = 0;
  for (int32_t v = 0; v < V; ++v) lp_all += deg(v) * (deg(v) - 1) / 2;
  int32_t best_h = 1;
  double best = -1;
  std::vector<unsigned long long> keys((size_t)E);
  for (int32_t h : {1, 2, 4, 6, 8, 10, 12, 16, 20, 24, 31}) {
    if (h > top) break;
    double lw = 0, lp = lp_all;
    for (int32_t i = 0; i < h; ++i) lp -= deg(order[i]) * (deg(order[i]) - 1) / 2;
    for (int32_t e = 0; e < E; ++e) {
      uint32_t m = 0;
      bool light = false;
      double sum = 0;
      for (uint32_t i = eoff[e]; i < eoff[e + 1]; ++i) {
        const int32_t v = enodes[i];
        if (idx[v] >= 0 && idx[v] < h) m |= 1u << idx[v];
        else {
          light = true;
          sum += deg(v) - 1;
        }
      }
      keys[e] = (unsigned long long)m << 1 | (light ? 1ull : 0ull);
      lw += sum * (sum - 1) / 2;
    }
    std::sort(keys.begin(), keys.end());
    const double K = (double)(std::unique(keys.begin(), keys.end()) - keys.begin());
    const double cost = K * K * K / 6 + lw + std::min(lp, 4 * K * K) * K;
    if (best < 0 || cost < best) {
      best = cost;
      best_h = h;
    }
  }
  *h_out = best_h;
  if (cost_out) *cost_out = best;
  return HM_OK;
}

template <typename MT>
static hm_status count_static_hub_impl(hm_graph* g, int32_t h, unsigned long long* h_acc) {
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  HubScratch& hb = g->hub;
  const int32_t E = c.E, V = c.V;
  hipStream_t st = g->stream;
  const size_t lds_centre = hub_centre_lds<MT>(E), lds_tri = hub_tri_lds<MT>(E);
  const size_t lds_fill = (size_t)E * sizeof(MT);
  HM_TRY(allow_lds((const void*)k_hub_centre<MT>, lds_centre));
  HM_TRY(allow_lds((const void*)k_hub_fill<MT>, lds_fill));
  HM_TRY(allow_lds((const void*)k_hub_tri<MT>, lds_tri));

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
  for (int32_t i = 0; i < hh; ++i) hubidx[order[i]] = (int8_t)i;
  unsigned long long lp_bound = 0;
  for (int32_t v = 0; v < V; ++v)
    if (hubidx[v] < 0) {
      const unsigned long long d = voff[v + 1] - voff[v];
      lp_bound += d * (d - (d > 0)) / 2;
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
  HM_TRY(hb.nruns.ensure(sizeof(unsigned long long) * 4));
  HM_TRY(hb.hmr.ensure(sizeof(uint32_t) * E));
  HM_TRY(hb.lszr.ensure((size_t)E));
  HM_TRY(hb.types.ensure(sizeof(unsigned long long) * std::max<unsigned long long>(lp_bound, 1)));
  HM_TRY(hb.types2.ensure(sizeof(unsigned long long) * std::max<unsigned long long>(lp_bound, 1)));
  HM_TRY(hb.tcnt.ensure(sizeof(uint32_t) * std::max<unsigned long long>(lp_bound, 1)));
  HM_TRY(hb.tkey.ensure(sizeof(unsigned long long) * std::max<unsigned long long>(lp_bound, 1)));
  HM_TRY(s.deg.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.ids.ensure(sizeof(int32_t) * E));
  HM_TRY(s.perm.ensure(sizeof(int32_t) * E));
  HM_TRY(s.rank.ensure(sizeof(int32_t) * E));
  HM_TRY(s.degr.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.roff.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.rlen.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.work2.ensure(sizeof(unsigned long long) * E));
  HM_TRY(s.acc.ensure(sizeof(unsigned long long) * kAccCount));
  unsigned long long* acc = s.acc.as<unsigned long long>();
  unsigned long long* runs = hb.nruns.as<unsigned long long>();  // [0] classes, [1] light pairs, [2] types
  const unsigned eb = (unsigned)((E + kBlock - 1) / kBlock);
  const int tbits = 32 + hh + 1;
  const unsigned long long tcap = hb.types.bytes / sizeof(unsigned long long);

  HM_HIP(hipMemcpyAsync(hb.hubidx.p, hubidx.data(), (size_t)std::max(V, 1), hipMemcpyHostToDevice, st));
  HM_HIP(hipEventRecord(g->ev[0], st));
  k_hub_zero<<<1, 64, 0, st>>>(acc, kAccCount);
  k_hub_zero<<<1, 64, 0, st>>>(runs, 4);
  k_hub_class<<<eb, kBlock, 0, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), hb.hubidx.as<int8_t>(),
                                     hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(),
                                     hb.ckey.as<unsigned long long>());
  HM_HIP(hipGetLastError());

  // (A) classes: sorted unique keys with counts
  size_t tb1 = 0, tb2 = 0, tb3 = 0, tb4 = 0, tb5 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, hb.ckey.as<unsigned long long>(),
                                           hb.ckey2.as<unsigned long long>(), E, 0, hh + 1, st));
  HM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, hb.ckey2.as<unsigned long long>(),
                                               hb.ukey.as<unsigned long long>(), hb.ucnt.as<uint32_t>(),
                                               runs, E, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb5, hb.types.as<unsigned long long>(),
                                           hb.types2.as<unsigned long long>(), (int)std::min<unsigned long long>(tcap, INT32_MAX),
                                           0, tbits, st));
  size_t tbm = std::max(std::max(tb1, tb2), std::max(std::max(tb3, tb4), tb5));
  HM_TRY(s.tmp.ensure(tbm));
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
  if (K > k0) {
    k_hub_classes<<<(unsigned)(K - k0), kBlock, 0, st>>>(K, k0, hb.ukey.as<unsigned long long>(),
                                                         hb.ucnt.as<uint32_t>(), acc);
    HM_HIP(hipGetLastError());
  }
  HM_HIP(hipEventRecord(g->ev[1], st));

  // light neighbours: (C), light degrees, pair types of the light pairs for (B)
  k_hub_centre<MT><<<E, kBlock, lds_centre, st>>>(
      E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
      hb.hubidx.as<int8_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(), s.deg.as<uint32_t>(),
      hb.types.as<unsigned long long>(), runs + 1, tcap, acc);
  HM_HIP(hipGetLastError());
  unsigned long long nt = 0;
  HM_HIP(hipMemcpyAsync(&nt, runs + 1, sizeof(nt), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  if (nt > tcap) return fail(HM_ERR_DEVICE, "hub-aware count: more light pairs than their bound");
  g->hub_stats.light_pairs = (int64_t)nt;
  HM_HIP(hipEventRecord(g->ev[2], st));
  if (nt > 0) {
    tb = s.tmp.bytes;
    HM_HIP(hipcub::DeviceRadixSort::SortKeys(s.tmp.p, tb, hb.types.as<unsigned long long>(),
                                             hb.types2.as<unsigned long long>(), (int)nt, 0, tbits, st));
    size_t need = 0;
    HM_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, need, hb.types2.as<unsigned long long>(),
                                                 hb.tkey.as<unsigned long long>(), hb.tcnt.as<uint32_t>(),
                                                 runs + 2, (int)nt, st));
    HM_TRY(s.tmp.ensure(need));
    tb = s.tmp.bytes;
    HM_HIP(hipcub::DeviceRunLengthEncode::Encode(s.tmp.p, tb, hb.types2.as<unsigned long long>(),
                                                 hb.tkey.as<unsigned long long>(), hb.tcnt.as<uint32_t>(),
                                                 runs + 2, (int)nt, st));
    unsigned long long T = 0;
    HM_HIP(hipMemcpyAsync(&T, runs + 2, sizeof(T), hipMemcpyDeviceToHost, st));
    HM_HIP(hipStreamSynchronize(st));
    g->hub_stats.pair_types = (int64_t)T;
    if (T > 0 && K > k0) {
      k_hub_types<<<(unsigned)T, kBlock, 0, st>>>(K, k0, hb.ukey.as<unsigned long long>(),
                                                   hb.ucnt.as<uint32_t>(), hb.tkey.as<unsigned long long>(),
                                                   hb.tcnt.as<uint32_t>(), acc);
      HM_HIP(hipGetLastError());
    }
  }
  HM_HIP(hipEventRecord(g->ev[3], st));

  // (D) light triangles over the light projection, ranked by ascending light degree
  k_hub_iota<<<eb, kBlock, 0, st>>>(s.ids.as<int32_t>(), E);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, tb, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  k_hub_rank<<<eb, kBlock, 0, st>>>(E, s.perm.as<int32_t>(), hb.hmask.as<uint32_t>(), hb.lsz.as<uint8_t>(),
                                    s.rank.as<int32_t>(), hb.hmr.as<uint32_t>(), hb.lszr.as<uint8_t>());
  k_hub_widen<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      s.work2.as<uint32_t>(), s.degr.as<unsigned long long>(), E);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  unsigned long long total = 0;
  HM_HIP(hipMemcpyAsync(&total, s.roff.as<unsigned long long>() + E, sizeof(total), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  g->hub_stats.light_wedges = -1;
  HM_TRY(s.rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total, 1)));
  k_hub_fill<MT><<<E, kBlock, lds_fill, st>>>(
      E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
      hb.hubidx.as<int8_t>(), hb.lsz.as<uint8_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
      s.rlen.as<uint32_t>(), s.rows.as<uint32_t>());
  HM_HIP(hipGetLastError());
  k_hub_tri<MT><<<E, kBlock, lds_tri, st>>>(
      E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
      hb.hubidx.as<int8_t>(), s.rank.as<int32_t>(), s.perm.as<int32_t>(), hb.hmr.as<uint32_t>(),
      hb.lszr.as<uint8_t>(), s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(), acc);
  HM_HIP(hipGetLastError());
  HM_HIP(hipEventRecord(g->ev[4], st));
  HM_HIP(hipMemcpyAsync(h_acc, acc, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  // the projection statistics are not computed on this path
  h_acc[kAccP] = h_acc[kAccMaxDeg] = h_acc[kAccOrient] = h_acc[kAccWedges] = h_acc[kAccUpper] = ~0ull;
  unsigned long long dmax = 0;
  for (int32_t v = 0; v < V; ++v) dmax = std::max<unsigned long long>(dmax, voff[v + 1] - voff[v]);
  h_acc[kAccMaxNodeDeg] = dmax;
  float m01 = 0, m12 = 0, m23 = 0, m34 = 0;
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

hm_status count_static_hub(hm_graph* g, int32_t h, unsigned long long* h_acc) {
  if (h < 1 || h > kHubMax) return fail(HM_ERR_BAD_INPUT, "hub count outside 1..31");
  return g->csr.max_size > 32 ? count_static_hub_impl<mask64_t>(g, h, h_acc)
                              : count_static_hub_impl<uint32_t>(g, h, h_acc);
}

}  // namespace hm
