// Per-hyperedge bodies of the single-table static count, shared by the one-graph kernels
// (hm_static.hip) and the many-graph kernels (hm_batch.hip).
//
// Every body is run by one whole workgroup for one hyperedge of one graph; the pointers it gets
// are that graph's own arrays, indexed by the graph's local ids and ranks. The dynamic LDS holds
// the mask table M over the graph's ids or ranks (and, for the triangle body, the size table
// after it), so it must span the largest graph the kernel sees.
//
// Counting scheme (MoCHy-master/main_exact_par.cpp:74-173 counts every connected triple once):
//   open triples   — OpenAll(a): every unordered pair of neighbours of a classified as an open
//                    triple centred at a, from the histogram of neighbour classes (C_ab, [b ⊄ a]);
//   closed triples — each triangle a < b < c in a degree order adds +1 to its closed bin and -1 to
//                    the "as if open" bin of each of its three centres.
#pragma once
#include <hip/hip_runtime.h>

#include "hm_device.hpp"
#include "hm_tables.hpp"

namespace hm {

static __constant__ int8_t c_static_bin[128] = HM_STATIC_BIN_TABLE;


// OpenAll(a): every unordered pair of neighbours of a, classified as an open triple centred at a
// from the neighbour classes cls[2 (C_ab - 1) + [b ⊄ a]] (sa = |a|). Adds to acc bins 20..25.
__device__ inline void open_all_pairs(const uint32_t* cls, int sa, unsigned long long* acc) {
  __shared__ unsigned long long red[6][kWavesPerBlock];
  const int K = 2 * sa;
  unsigned long long o0 = 0, o1 = 0, o2 = 0, o3 = 0, o4 = 0, o5 = 0;
  for (int p = threadIdx.x; p < K * K; p += blockDim.x) {
    const int k1 = p / K, k2 = p - k1 * K;
    if (k1 > k2) continue;
    const unsigned long long n1 = cls[k1], n2 = cls[k2];
    const unsigned long long pairs = (k1 == k2) ? n1 * (n1 - (n1 > 0)) / 2 : n1 * n2;
    if (!pairs) continue;
    const int c1 = (k1 >> 1) + 1, c2 = (k2 >> 1) + 1;
    const int b6 = (int)(c1 + c2 < sa) + 2 * ((k1 & 1) + (k2 & 1));
    o0 += b6 == 0 ? pairs : 0;
    o1 += b6 == 1 ? pairs : 0;
    o2 += b6 == 2 ? pairs : 0;
    o3 += b6 == 3 ? pairs : 0;
    o4 += b6 == 4 ? pairs : 0;
    o5 += b6 == 5 ? pairs : 0;
  }
  o0 = wave_sum(o0); o1 = wave_sum(o1); o2 = wave_sum(o2);
  o3 = wave_sum(o3); o4 = wave_sum(o4); o5 = wave_sum(o5);
  if (lane_id() == 0) {
    const int w = wave_id();
    red[0][w] = o0; red[1][w] = o1; red[2][w] = o2;
    red[3][w] = o3; red[4][w] = o4; red[5][w] = o5;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += red[threadIdx.x][w];
    if (t) atomicAdd(&acc[kAccHist + kOpenBin0 + threadIdx.x], t);
  }
}

// Hyperedge a (kBlock lanes): projected degree *deg_out, neighbours with a smaller id *dlow_out
// (if not null) and OpenAll(a) into acc. M spans E ids.
template <typename MT>
__device__ inline void proj_count_item(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                       const int32_t* __restrict__ enodes,
                                       const uint8_t* __restrict__ esize,
                                       const uint32_t* __restrict__ voff,
                                       const int32_t* __restrict__ vedges, uint32_t* deg_out,
                                       uint32_t* dlow_out, unsigned long long* acc) {
  __shared__ uint32_t cls[2 * HM_MAX_EDGE_SIZE];
  __shared__ uint32_t dred[2][kWavesPerBlock];
  for (int k = threadIdx.x; k < 2 * HM_MAX_EDGE_SIZE; k += blockDim.x) cls[k] = 0;
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, nullptr);

  // neighbour classes: k = 2 (C_ab - 1) + [|b| > C_ab]
  uint32_t d = 0, dl = 0;
  for (int x = threadIdx.x; x < E; x += blockDim.x) {
    const MT m = M[x];
    if (m) {
      const int C = popc(m);
      atomicAdd(&cls[2 * (C - 1) + (esize[x] > C ? 1 : 0)], 1u);
      ++d;
      dl += x < a ? 1u : 0u;
    }
  }
  d = wave_sum(d);
  dl = wave_sum(dl);
  if (lane_id() == 0) {
    dred[0][wave_id()] = d;
    dred[1][wave_id()] = dl;
  }
  __syncthreads();

  open_all_pairs(cls, (int)(eoff[a + 1] - eoff[a]), acc);
  if (threadIdx.x == 0) {
    uint32_t dd = 0, ll = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
      dd += dred[0][w];
      ll += dred[1][w];
    }
    *deg_out = dd;
    if (dlow_out) *dlow_out = ll;
  }
}

// Hyperedge a (kBlock lanes): its upper row — ranks r > rank[a] with M[r] != 0, ascending, as
// packed (r << 7 | C_ar) words at rows[roff[rank[a]]...] — and its length rlen[rank[a]]. M spans
// E ranks. With `hub` given, only a's nodes v with hub[v] < 0 take part (the light projection).
template <typename MT>
__device__ inline void proj_fill_item(MT* M, int32_t E, int32_t a, const uint32_t* __restrict__ eoff,
                                      const int32_t* __restrict__ enodes,
                                      const uint32_t* __restrict__ voff,
                                      const int32_t* __restrict__ vedges,
                                      const int32_t* __restrict__ rank,
                                      const unsigned long long* __restrict__ roff,
                                      uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows,
                                      const int8_t* __restrict__ hub = nullptr) {
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t ra = rank[a];
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, rank, hub);

  const unsigned long long base = roff[ra];
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
      rows[base + pos] = ((uint32_t)r << kCntBits) | (uint32_t)popc(m);
    }
    running += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) rlen[ra] = running;
}

constexpr int kTriLoads = 4;

// Slots of a closed triple (a, b, c): its bin and the three bins it has as if open at a, at b and
// at c follow from the Venn id and, per private region x_only (x_only = s_x - C_xy - C_xz + g),
// the bit [x_only > g] = [s_x > C_xy + C_xz] that decides whether x keeps a private node once the
// triple region is counted as if it were shared by y and z only. slot = pairs * 27 + t_a + 3 t_b
// + 9 t_c with t_x = [x_only > 0] + [x_only > g] and pairs = [C_ab > g] | [C_bc > g] << 1 |
// [C_ca > g] << 2 (slot group 8: g = 0, every pair region non-empty, t_x in {0, 2}).
constexpr int kTriSlots = 243;

// Closed bin's Venn id and the three as-if-open bins (centres a, b, c) of slot p.
__device__ inline void tri_slot_bins(int p, int* v, int* oa, int* ob, int* oc) {
  const int grp = p / 27, t = p - grp * 27;
  const int ta = t % 3, tb = (t / 3) % 3, tc = t / 9;
  const int pairs = grp < 8 ? grp : 7;
  const int b0 = ta > 0, b1 = tb > 0, b2 = tc > 0;
  const int ab = pairs & 1, bc = (pairs >> 1) & 1, ca = (pairs >> 2) & 1;
  *v = b0 | b1 << 1 | b2 << 2 | pairs << 3 | (grp < 8 ? 64 : 0);
  // open_bin(s_x, s_y, s_z, C_xy, C_xz) = kOpenBin0 + [x_only > g] + 2 ([y ⊄ x] + [z ⊄ x]), with
  // [b ⊄ a] = [b_only + bc_only > 0] and likewise for the other five containments
  *oa = kOpenBin0 + (ta == 2) + 2 * ((b1 | bc) + (b2 | bc));
  *ob = kOpenBin0 + (tb == 2) + 2 * ((b0 | ca) + (b2 | ca));
  *oc = kOpenBin0 + (tc == 2) + 2 * ((b0 | ab) + (b1 | ab));
}

// Slot of a closed triple, from sign bits ([x < 0] = (unsigned)x >> 31, all terms small) rather
// than compare-and-select chains: ka = C_ab - s_a and kb = C_ab - s_b are fixed per b, so
//   [a_only > 0] = [ka + C_ac - g < 0],   [a_only > g] = [ka + C_ac < 0],
//   [b_only > 0] = [kb + C_bc - g < 0],   [b_only > g] = [kb + C_bc < 0],
//   [c_only > 0] = [C_ac - g + C_bc - s_c < 0],   [c_only > g] = [C_ac + C_bc - s_c < 0],
// the pair bits [C_xy > g] = [g - C_xy < 0], and group 8 (g = 0, pairs 7) = pairs + [g - 1 < 0].
__device__ inline unsigned tri_slot(int ka, int kb, int cab, int sc, int cbc, int cac, int g) {
  const int d = cac - g, e = cbc - g;
  const unsigned ta = ((unsigned)(ka + d) >> 31) + ((unsigned)(ka + cac) >> 31);
  const unsigned tb = ((unsigned)(kb + e) >> 31) + ((unsigned)(kb + cbc) >> 31);
  const unsigned tc = ((unsigned)(d + cbc - sc) >> 31) + ((unsigned)(cac + cbc - sc) >> 31);
  const unsigned grp = ((unsigned)(g - cab) >> 31) + ((unsigned)(-e) >> 31 << 1) +
                       ((unsigned)(-d) >> 31 << 2) + ((unsigned)(g - 1) >> 31);
  return grp * 27u + ta + 3u * tb + 9u * tc;
}

// Centre rank r (NW waves; NC copies of the slot counters shared by the workgroup, copy = lane %
// NC): every triangle r < rb < rc with rb in r's upper row and rc in rb's upper row, one LDS
// atomic each into its slot; the slots are folded into the 30 bins once per centre.
// M (mask table by rank) spans E ranks; the E size bytes follow it in the same dynamic LDS.
template <typename MT, int NW, int NC>
__device__ inline void tri_item(MT* M, int32_t E, int32_t r, const uint32_t* __restrict__ eoff,
                                const int32_t* __restrict__ enodes,
                                const uint32_t* __restrict__ voff,
                                const int32_t* __restrict__ vedges,
                                const int32_t* __restrict__ rank, const int32_t* __restrict__ perm,
                                const uint8_t* __restrict__ szr,
                                const unsigned long long* __restrict__ roff,
                                const uint32_t* __restrict__ rlen, const uint32_t* __restrict__ rows,
                                unsigned long long* __restrict__ acc) {
  // the hosts reserve 20 KB of LDS beside the dynamic table for these counters
  static_assert(kTriSlots * NC * sizeof(uint32_t) + 32 * sizeof(unsigned long long) <= 16 * 1024,
                "slot counters outgrow the static LDS reservation");
  __shared__ uint32_t hs[kTriSlots * NC];
  __shared__ unsigned long long hb[32];
  uint8_t* sz = reinterpret_cast<uint8_t*>(M + E);
  const uint32_t nA = rlen[r];
  if (nA < 2) return;  // no triangle has r as its lowest-ranked member
  const int32_t a = perm[r];
  for (int k = threadIdx.x; k < kTriSlots * NC; k += blockDim.x) hs[k] = 0u;
  if (threadIdx.x < 32) hb[threadIdx.x] = 0ull;
  for (int x = threadIdx.x; x < E; x += blockDim.x) sz[x] = szr[x];
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, rank);

  const int lane = lane_id(), w = wave_id();
  uint32_t* hc = hs + (lane & (NC - 1));
  const int sa = sz[r];
  const uint32_t* rowA = rows + roff[r];
  for (uint32_t ib = w; ib < nA; ib += NW) {
    const uint32_t eb = rowA[ib];
    const int32_t rb = (int32_t)(eb >> kCntBits);
    const int cab = (int)(eb & kCntMask);
    const MT mab = M[rb];
    const int ka = cab - sa, kb = cab - sz[rb];
    const uint32_t* rowB = rows + roff[rb];
    const uint32_t nB = rlen[rb];
    auto visit = [&](uint32_t ec) {
      const int32_t rc = (int32_t)(ec >> kCntBits);
      const MT mc = M[rc];
      if (mc) {
        const int cbc = (int)(ec & kCntMask);
        const int cac = popc(mc);
        const int g = popc((MT)(mab & mc));
        const int sc = sz[rc];
        atomicAdd(&hc[tri_slot(ka, kb, cab, sc, cbc, cac, g) * NC], 1u);
      }
    };
    // kTriLoads row loads in flight per lane before any is used (lane l reads l, l + 64, ...)
    uint32_t j = lane;
    for (; j + (kTriLoads - 1) * kWave < nB; j += kTriLoads * kWave) {
      uint32_t e[kTriLoads];
#pragma unroll
      for (int u = 0; u < kTriLoads; ++u) e[u] = rowB[j + u * kWave];
#pragma unroll
      for (int u = 0; u < kTriLoads; ++u) visit(e[u]);
    }
    for (; j < nB; j += kWave) visit(rowB[j]);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < kTriSlots; p += blockDim.x) {
    unsigned long long n = 0;
    for (int c = 0; c < NC; ++c) n += hs[p * NC + c];
    if (n) {
      int v, oa, ob, oc;
      tri_slot_bins(p, &v, &oa, &ob, &oc);
      atomicAdd(&hb[c_static_bin[v]], n);
      atomicAdd(&hb[oa], 0ull - n);
      atomicAdd(&hb[ob], 0ull - n);
      atomicAdd(&hb[oc], 0ull - n);
    }
  }
  __syncthreads();
  if (threadIdx.x < kStaticBins && hb[threadIdx.x]) atomicAdd(&acc[kAccHist + threadIdx.x], hb[threadIdx.x]);
}

}  // namespace hm
