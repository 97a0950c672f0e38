// Static exact h-motif counting on MI355X (gfx950).
//
// Replaces dynamicHyperGraph/src/HMotifCount.cu:16-125 (motifTriangleKernel / computeMotifCounts)
// and follows the counting semantics of MoCHy-E (MoCHy-master/main_exact_par.cpp:74-173):
// every connected triple of distinct hyperedges is counted once, into one of 30 bins.
//
// Decomposition (all integer, exact):
//   open triples   — counted combinatorially per centre a: the number of pairs {b, c} of
//                    neighbours of a per open bin follows from the histogram of neighbour classes
//                    (C_ab, [b ⊄ a]); this counts every pair as if it were open.
//   closed triples — enumerated once each as oriented triangles a < b < c in a degree order
//                    (rank = position in ascending projected degree): +1 in the closed bin and
//                    -1 in the "as if open" bin of each of its three centres, which cancels their
//                    share of the combinatorial open count.
// So  H = sum_a OpenAll(a) + sum_{triangles} ( e_closed - e_open(a) - e_open(b) - e_open(c) ).
//
// Kernels (one workgroup of 256 lanes per hyperedge):
//   k_proj_count  LDS mask table M[x] = bitmask of a's nodes contained in x, for every x (dense
//                 over the hyperedge ids); projected degree d_a and the open-pair classes.
//   k_proj_fill   same table indexed by rank; writes a's upper row (rank > rank(a)) as packed
//                 (rank << 7 | C) words in ascending rank.
//   k_tri         same table by rank; for every b in a's upper row, scans b's upper row:
//                 M[c] != 0 <=> triangle, C_ac = popc(M[c]), g_abc = popc(M[b] & M[c]).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_tables.hpp"

namespace hm {

__constant__ int8_t c_static_bin[128] = HM_STATIC_BIN_TABLE;

// Accumulator slots in CountScratch::acc (uint64).
enum AccSlot : int {
  kAccHist = 0,        // 30 bins (two's-complement wrap: intermediate -1 contributions are exact)
  kAccP = 32,          // sum_a d_a
  kAccMaxDeg = 33,     // max_a d_a
  kAccOrient = 34,     // sum_a sum_{b in N(a), b > a} d_b
  kAccWedges = 35,     // sum_a C(d_a, 2)
  kAccUpper = 36,      // sum_a d+_a (upper-row entries)
  kAccMaxNodeDeg = 37, // max_v |V2H(v)|
  kAccCount = 40
};

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_count(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, uint32_t* __restrict__ deg,
    unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t cls[2 * HM_MAX_EDGE_SIZE];
  __shared__ unsigned long long red[6][kWavesPerBlock];
  __shared__ uint32_t dred[kWavesPerBlock];
  const int32_t a = blockIdx.x;
  for (int k = threadIdx.x; k < 2 * HM_MAX_EDGE_SIZE; k += blockDim.x) cls[k] = 0;
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, nullptr);

  // neighbour classes: k = 2 (C_ab - 1) + [|b| > C_ab]
  uint32_t d = 0;
  for (int x = threadIdx.x; x < E; x += blockDim.x) {
    const MT m = M[x];
    if (m) {
      const int C = popc(m);
      atomicAdd(&cls[2 * (C - 1) + (esize[x] > C ? 1 : 0)], 1u);
      ++d;
    }
  }
  d = wave_sum(d);
  if (lane_id() == 0) dred[wave_id()] = d;
  __syncthreads();

  // OpenAll(a): every unordered pair of neighbours, classified as an open triple centred at a.
  const int sa = (int)(eoff[a + 1] - eoff[a]);
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
    unsigned long long s = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) s += red[threadIdx.x][w];
    if (s) atomicAdd(&acc[kAccHist + kOpenBin0 + threadIdx.x], s);
  }
  if (threadIdx.x == 0) {
    uint32_t dd = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) dd += dred[w];
    deg[a] = dd;
  }
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_fill(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ perm,
    const uint32_t* __restrict__ deg, const unsigned long long* __restrict__ roff,
    uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows,
    unsigned long long* __restrict__ orient) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t wcnt[kWavesPerBlock];
  __shared__ unsigned long long ored[kWavesPerBlock];
  const int32_t a = blockIdx.x;
  const int32_t ra = rank[a];
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, rank);

  // compact ranks r > ra with M[r] != 0, ascending
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

  // statistic for the roofline model: sum of d_b over neighbours b with id > a
  unsigned long long s = 0;
  for (int32_t r = threadIdx.x; r < E; r += blockDim.x) {
    if (M[r]) {
      const int32_t b = perm[r];
      if (b > a) s += deg[b];
    }
  }
  s = wave_sum(s);
  if (lane == 0) ored[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) t += ored[q];
    orient[a] = t;
  }
}

// Work estimate of centre r for the triangle kernel: sum of upper-row lengths of its upper row.
__global__ __launch_bounds__(kBlock) void k_work(int32_t E, const unsigned long long* __restrict__ roff,
                                                 const uint32_t* __restrict__ rlen,
                                                 const uint32_t* __restrict__ rows,
                                                 unsigned long long* __restrict__ work) {
  const int32_t r = blockIdx.x * kWavesPerBlock + wave_id();
  if (r >= E) return;
  const unsigned long long s0 = roff[r];
  const uint32_t n = rlen[r];
  unsigned long long s = 0;
  for (uint32_t j = lane_id(); j < n; j += kWave) s += rlen[rows[s0 + j] >> kCntBits];
  s = wave_sum(s);
  if (lane_id() == 0) work[r] = s + n;
}

// 8 waves per centre workgroup (c1: 4 waves 9.03 ms, 8 waves 7.85 ms, 16 waves 8.09 ms).
constexpr int kTriWaves = 8;

template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_tri(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ perm,
    const uint8_t* __restrict__ szr, const unsigned long long* __restrict__ roff,
    const uint32_t* __restrict__ rlen, const uint32_t* __restrict__ rows,
    const int32_t* __restrict__ order, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  uint8_t* sz = reinterpret_cast<uint8_t*>(M + E);
  constexpr int kCopies = 8;
  __shared__ int32_t h[NW][32][kCopies];
  __shared__ int8_t tbl[128];
  const int32_t r = order[blockIdx.x];
  const uint32_t nA = rlen[r];
  if (nA < 2) return;  // no triangle has r as its lowest-ranked member
  const int32_t a = perm[r];
  for (int k = threadIdx.x; k < NW * 32 * kCopies; k += blockDim.x)
    (&h[0][0][0])[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_static_bin[threadIdx.x];
  for (int x = threadIdx.x; x < E; x += blockDim.x) sz[x] = szr[x];
  build_masks<MT>(M, E, a, eoff, enodes, voff, vedges, rank);

  const int lane = lane_id(), w = wave_id();
  const int copy = lane & (kCopies - 1);
  const int sa = sz[r];
  const uint32_t* rowA = rows + roff[r];
  for (uint32_t ib = w; ib < nA; ib += NW) {
    const uint32_t eb = rowA[ib];
    const int32_t rb = (int32_t)(eb >> kCntBits);
    const int cab = (int)(eb & kCntMask);
    const MT mab = M[rb];
    const int sb = sz[rb];
    const uint32_t* rowB = rows + roff[rb];
    const uint32_t nB = rlen[rb];
    int32_t* hc = &h[w][0][copy];
    auto visit = [&](uint32_t ec) {
      const int32_t rc = (int32_t)(ec >> kCntBits);
      const MT mc = M[rc];
      if (mc) {
        const int cbc = (int)(ec & kCntMask);
        const int cac = popc(mc);
        const int g = popc((MT)(mab & mc));
        const int sc = sz[rc];
        const int closed = tbl[venn_id(sa, sb, sc, cab, cbc, cac, g)];
        atomicAdd(&hc[closed * kCopies], 1);
        atomicSub(&hc[open_bin(sa, sb, sc, cab, cac) * kCopies], 1);
        atomicSub(&hc[open_bin(sb, sa, sc, cab, cbc) * kCopies], 1);
        atomicSub(&hc[open_bin(sc, sa, sb, cac, cbc) * kCopies], 1);
      }
    };
    // two row loads in flight per lane before either is used (lane l reads l, l + 64, ...)
    uint32_t j = lane;
    for (; j + kWave < nB; j += 2 * kWave) {
      const uint32_t e0 = rowB[j], e1 = rowB[j + kWave];
      visit(e0);
      visit(e1);
    }
    if (j < nB) visit(rowB[j]);
  }
  __syncthreads();
  if (threadIdx.x < kStaticBins) {
    long long s = 0;
    for (int q = 0; q < NW; ++q)
      for (int c = 0; c < kCopies; ++c) s += h[q][threadIdx.x][c];
    if (s) atomicAdd(&acc[kAccHist + threadIdx.x], (unsigned long long)s);
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// out[i] = in[i] for i < n, out[n] = 0 (input of an exclusive scan with n + 1 outputs).
__global__ void k_widen(const uint32_t* __restrict__ in, unsigned long long* __restrict__ out,
                        int32_t n) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
  else if (i == n) out[n] = 0;
}

__global__ void k_iota(int32_t* p, int32_t n) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void k_rank_from_perm(const int32_t* __restrict__ perm, const uint8_t* __restrict__ esize,
                                 int32_t* __restrict__ rank, uint8_t* __restrict__ szr, int32_t n) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) {
    const int32_t a = perm[r];
    rank[a] = r;
    szr[r] = esize[a];
  }
}

// Graph statistics: P, max degree, wedges, upper entries, oriented row sum.
__global__ __launch_bounds__(kBlock) void k_stats(int32_t E, int32_t V, const uint32_t* __restrict__ voff,
                                                  const uint32_t* __restrict__ deg,
                                                  const unsigned long long* __restrict__ orient,
                                                  const uint32_t* __restrict__ rlen,
                                                  unsigned long long* __restrict__ acc) {
  unsigned long long p = 0, mx = 0, wd = 0, up = 0, orr = 0;
  for (int32_t a = blockIdx.x * blockDim.x + threadIdx.x; a < E; a += gridDim.x * blockDim.x) {
    const unsigned long long d = deg[a];
    p += d;
    mx = d > mx ? d : mx;
    wd += d * (d - (d > 0)) / 2;
    up += rlen[a];
    orr += orient[a];
  }
  unsigned long long mv = 0;
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
    mv = max(mv, (unsigned long long)(voff[v + 1] - voff[v]));
  p = wave_sum(p); wd = wave_sum(wd); up = wave_sum(up); orr = wave_sum(orr); mx = wave_max(mx);
  mv = wave_max(mv);
  if (lane_id() == 0) {
    atomicAdd(&acc[kAccP], p);
    atomicMax(&acc[kAccMaxDeg], mx);
    atomicAdd(&acc[kAccWedges], wd);
    atomicAdd(&acc[kAccUpper], up);
    atomicAdd(&acc[kAccOrient], orr);
    atomicMax(&acc[kAccMaxNodeDeg], mv);
  }
}

// ---------------------------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------------------------
static int lds_limit_bytes(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    v = 65536;
  return v;
}

// Dynamic LDS of the dense kernels: masks (and sizes for k_tri), with 16-byte alignment.
template <typename MT>
static size_t dense_lds_bytes(int32_t E, bool with_sizes) {
  size_t b = (size_t)E * sizeof(MT);
  if (with_sizes) b += (size_t)E;
  return (b + 15) & ~(size_t)15;
}

template <typename MT>
static hm_status set_lds_attr(const void* fn, size_t bytes) {
  return allow_lds(fn, bytes);
}

template <typename MT>
static hm_status count_static_dense(hm_graph* g, unsigned long long* h_acc) {
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  const int32_t E = c.E;
  hipStream_t st = g->stream;
  const size_t lds_cnt = dense_lds_bytes<MT>(E, false);
  const size_t lds_tri = dense_lds_bytes<MT>(E, true);
  HM_TRY(set_lds_attr<MT>((const void*)k_proj_count<MT>, lds_cnt));
  HM_TRY(set_lds_attr<MT>((const void*)k_proj_fill<MT>, lds_cnt));
  HM_TRY(set_lds_attr<MT>((const void*)k_tri<MT, kTriWaves>, lds_tri));

  HM_TRY(s.deg.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.ids.ensure(sizeof(int32_t) * E));
  HM_TRY(s.perm.ensure(sizeof(int32_t) * E));
  HM_TRY(s.rank.ensure(sizeof(int32_t) * E));
  HM_TRY(s.degr.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.roff.ensure(sizeof(unsigned long long) * (E + 1)));
  HM_TRY(s.rlen.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.szr.ensure((size_t)E));
  HM_TRY(s.work.ensure(sizeof(unsigned long long) * E));
  HM_TRY(s.work2.ensure(sizeof(unsigned long long) * E));
  HM_TRY(s.order.ensure(sizeof(int32_t) * E));
  HM_TRY(s.ranks.ensure(sizeof(int32_t) * E));
  HM_TRY(s.orient.ensure(sizeof(unsigned long long) * E));
  HM_TRY(s.acc.ensure(sizeof(unsigned long long) * kAccCount));

  unsigned long long* acc = s.acc.as<unsigned long long>();
  const unsigned eb = (unsigned)((E + kBlock - 1) / kBlock);

  HM_HIP(hipEventRecord(g->ev[0], st));
  k_zero_u64<<<1, 64, 0, st>>>(acc, kAccCount);
  // 1. projected degrees + combinatorial open counts
  k_proj_count<MT><<<E, kBlock, lds_cnt, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                                c.esize.as<uint8_t>(), c.voff.as<uint32_t>(),
                                                c.vedges.as<int32_t>(), s.deg.as<uint32_t>(), acc);
  HM_HIP(hipGetLastError());
  // 2. degree order: stable radix sort of (d_a, a)
  k_iota<<<eb, kBlock, 0, st>>>(s.ids.as<int32_t>(), E);
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb3, s.work.as<unsigned long long>(),
                                                      s.work2.as<unsigned long long>(),
                                                      s.ranks.as<int32_t>(), s.order.as<int32_t>(),
                                                      E, 0, 64, st));
  HM_TRY(s.tmp.ensure(std::max(tb1, std::max(tb2, tb3))));
  size_t tb = s.tmp.bytes;
  // sorted keys land in work2 (as uint32), values in perm
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(s.tmp.p, tb, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  k_rank_from_perm<<<eb, kBlock, 0, st>>>(s.perm.as<int32_t>(), c.esize.as<uint8_t>(),
                                          s.rank.as<int32_t>(), s.szr.as<uint8_t>(), E);
  // 3. row capacity = d_a (by rank), offsets by exclusive scan
  k_widen<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      s.work2.as<uint32_t>(), s.degr.as<unsigned long long>(), E);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  unsigned long long total_cap = 0;
  HM_HIP(hipMemcpyAsync(&total_cap, s.roff.as<unsigned long long>() + E, sizeof(total_cap),
                        hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_TRY(s.rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total_cap, 1)));
  // 4. upper rows
  k_proj_fill<MT><<<E, kBlock, lds_cnt, st>>>(
      E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(),
      c.vedges.as<int32_t>(), s.rank.as<int32_t>(), s.perm.as<int32_t>(), s.deg.as<uint32_t>(),
      s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),
      s.orient.as<unsigned long long>());
  HM_HIP(hipGetLastError());
  // 5. centre order by work, most first
  k_work<<<(unsigned)((E + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, st>>>(
      E, s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),
      s.work.as<unsigned long long>());
  k_iota<<<eb, kBlock, 0, st>>>(s.ranks.as<int32_t>(), E);
  tb = s.tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(s.tmp.p, tb, s.work.as<unsigned long long>(),
                                                      s.work2.as<unsigned long long>(),
                                                      s.ranks.as<int32_t>(), s.order.as<int32_t>(),
                                                      E, 0, 64, st));
  HM_HIP(hipEventRecord(g->ev[1], st));
  // 6. triangles
#define HM_TRI_ARGS                                                                           \
  E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(),                        \
      c.vedges.as<int32_t>(), s.rank.as<int32_t>(), s.perm.as<int32_t>(), s.szr.as<uint8_t>(),   \
      s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),             \
      s.order.as<int32_t>(), acc
  k_tri<MT, kTriWaves><<<E, kTriWaves * kWave, lds_tri, st>>>(HM_TRI_ARGS);
#undef HM_TRI_ARGS
  HM_HIP(hipGetLastError());
  HM_HIP(hipEventRecord(g->ev[2], st));
  k_stats<<<64, kBlock, 0, st>>>(E, c.V, c.voff.as<uint32_t>(), s.deg.as<uint32_t>(), s.orient.as<unsigned long long>(),
                                 s.rlen.as<uint32_t>(), acc);
  HM_HIP(hipMemcpyAsync(h_acc, acc, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_HIP(hipEventElapsedTime(&g->ms_project, g->ev[0], g->ev[1]));
  HM_HIP(hipEventElapsedTime(&g->ms_triangles, g->ev[1], g->ev[2]));
  return HM_OK;
}

// ---------------------------------------------------------------------------------------------
// Host-side CSR preparation: rows sorted and de-duplicated, V2H by counting sort.
// ---------------------------------------------------------------------------------------------
hm_status prepare_host_rows(const int64_t* offsets, const int32_t* nodes, int32_t E,
                            std::vector<int64_t>* off, std::vector<int32_t>* out,
                            int32_t* vmax_out) {
  if (E < 0) return fail(HM_ERR_BAD_INPUT, "num_edges < 0");
  if (E > 0 && (!offsets || (!nodes && offsets[E] > 0)))
    return fail(HM_ERR_BAD_INPUT, "null offsets/nodes");
  if (E > 0 && offsets[0] != 0) return fail(HM_ERR_BAD_INPUT, "offsets[0] must be 0");
  off->assign((size_t)E + 1, 0);
  out->clear();
  if (E > 0) out->reserve((size_t)offsets[E]);
  int32_t vmax = -1;
  for (int32_t e = 0; e < E; ++e) {
    if (offsets[e + 1] < offsets[e]) return fail(HM_ERR_BAD_INPUT, "offsets must be non-decreasing");
    const size_t s = out->size();
    for (int64_t i = offsets[e]; i < offsets[e + 1]; ++i) {
      if (nodes[i] < 0) return fail(HM_ERR_BAD_INPUT, "negative node id");
      out->push_back(nodes[i]);
    }
    std::sort(out->begin() + (long)s, out->end());
    out->erase(std::unique(out->begin() + (long)s, out->end()), out->end());
    for (size_t i = s; i < out->size(); ++i) vmax = std::max(vmax, (*out)[i]);
    (*off)[(size_t)e + 1] = (int64_t)out->size();
  }
  *vmax_out = vmax;
  return HM_OK;
}

hm_status upload_csr(hm_graph* g, const std::vector<int64_t>& h_off,
                     const std::vector<int32_t>& h_nodes) {
  DeviceCsr& c = g->csr;
  const int32_t E = c.E, V = c.V;
  const int64_t I = (int64_t)h_nodes.size();
  if (I >= ((int64_t)1 << 32)) return fail(HM_ERR_OVER_BUDGET, "more than 2^32 incidences");
  c.I = I;
  std::vector<uint32_t> eoff((size_t)E + 1), voff((size_t)V + 1, 0);
  std::vector<uint8_t> esz((size_t)std::max(E, 1), 0);
  for (int32_t e = 0; e <= E; ++e) eoff[(size_t)e] = (uint32_t)h_off[(size_t)e];
  for (int32_t e = 0; e < E; ++e) {
    int64_t sz = h_off[(size_t)e + 1] - h_off[(size_t)e];
    esz[(size_t)e] = (uint8_t)std::min<int64_t>(sz, 255);
  }
  c.max_size = 0;
  for (int32_t e = 0; e < E; ++e)
    c.max_size = std::max<int32_t>(c.max_size, (int32_t)(h_off[(size_t)e + 1] - h_off[(size_t)e]));
  for (int64_t i = 0; i < I; ++i) voff[(size_t)h_nodes[(size_t)i] + 1]++;
  for (int32_t v = 0; v < V; ++v) voff[(size_t)v + 1] += voff[(size_t)v];
  std::vector<int32_t> vedges((size_t)std::max<int64_t>(I, 1));
  {
    std::vector<uint32_t> fill(voff.begin(), voff.end() - 1);
    for (int32_t e = 0; e < E; ++e)
      for (int64_t i = h_off[(size_t)e]; i < h_off[(size_t)e + 1]; ++i)
        vedges[fill[(size_t)h_nodes[(size_t)i]]++] = e;
  }
  HM_TRY(c.eoff.ensure(sizeof(uint32_t) * (E + 1)));
  HM_TRY(c.enodes.ensure(sizeof(int32_t) * std::max<int64_t>(I, 1)));
  HM_TRY(c.esize.ensure((size_t)std::max(E, 1)));
  HM_TRY(c.voff.ensure(sizeof(uint32_t) * ((size_t)V + 1)));
  HM_TRY(c.vedges.ensure(sizeof(int32_t) * std::max<int64_t>(I, 1)));
  HM_HIP(hipMemcpyAsync(c.eoff.p, eoff.data(), sizeof(uint32_t) * (E + 1), hipMemcpyHostToDevice, g->stream));
  if (I > 0) {
    HM_HIP(hipMemcpyAsync(c.enodes.p, h_nodes.data(), sizeof(int32_t) * I, hipMemcpyHostToDevice, g->stream));
    HM_HIP(hipMemcpyAsync(c.vedges.p, vedges.data(), sizeof(int32_t) * I, hipMemcpyHostToDevice, g->stream));
  }
  HM_HIP(hipMemcpyAsync(c.esize.p, esz.data(), (size_t)std::max(E, 1), hipMemcpyHostToDevice, g->stream));
  HM_HIP(hipMemcpyAsync(c.voff.p, voff.data(), sizeof(uint32_t) * ((size_t)V + 1), hipMemcpyHostToDevice, g->stream));
  HM_HIP(hipStreamSynchronize(g->stream));
  return HM_OK;
}

hm_status count_static(hm_graph* g, uint64_t out[HM_STATIC_BINS]) {
  const int32_t E = g->csr.E;
  unsigned long long acc[kAccCount] = {0};
  g->ms_project = g->ms_triangles = 0.f;
  if (E >= 3) {
    if (g->csr.max_size > HM_MAX_EDGE_SIZE)
      return fail(HM_ERR_UNSUPPORTED, "hyperedge larger than 64 nodes");
    if ((int64_t)E >= kMaxPackedId) return fail(HM_ERR_UNSUPPORTED, "more than 2^25 hyperedges");
    HM_HIP(hipSetDevice(g->device));
    const int lim = lds_limit_bytes(g->device);
    if (g->csr.max_size <= 32 && dense_lds_bytes<uint32_t>(E, true) <= (size_t)lim) {
      HM_TRY(count_static_dense<uint32_t>(g, acc));
    } else if (dense_lds_bytes<mask64_t>(E, true) <= (size_t)lim) {
      HM_TRY(count_static_dense<mask64_t>(g, acc));
    } else {
      set_error("static count: %d hyperedges exceed the dense LDS table (%d bytes per workgroup)",
                E, lim);
      return HM_ERR_UNSUPPORTED;
    }
  }
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = acc[kAccHist + k];
  hm_graph_stats& st = g->stats;
  st.num_edges = E;
  st.num_nodes = g->csr.V;
  st.num_incidences = g->csr.I;
  st.proj_pairs = (int64_t)acc[kAccP];
  st.max_proj_degree = (int64_t)acc[kAccMaxDeg];
  st.oriented_row_sum = (int64_t)acc[kAccOrient];
  st.wedges = (int64_t)acc[kAccWedges];
  int64_t closed = 0;
  for (int k = 0; k < HM_STATIC_BINS; ++k)
    if (k < kOpenBin0 || k >= kOpenBin0 + 6) closed += (int64_t)out[k];
  st.triangles = closed;
  st.max_node_degree = (int64_t)acc[kAccMaxNodeDeg];
  g->have_stats = true;
  return HM_OK;
}

}  // namespace hm

extern "C" {

hm_status hm_graph_create(const int64_t* offsets, const int32_t* nodes, int32_t num_edges,
                          int32_t num_nodes, hm_graph** out) {
  hm::clear_error();
  if (!out) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_create: null out");
  *out = nullptr;
  hm_graph* g = new (std::nothrow) hm_graph();
  if (!g) return hm::fail(HM_ERR_OOM, "hm_graph_create: host allocation failed");
  int32_t vmax = -1;
  std::vector<int64_t> h_off;
  std::vector<int32_t> h_nodes;
  hm_status st = hm::prepare_host_rows(offsets, nodes, num_edges, &h_off, &h_nodes, &vmax);
  if (st != HM_OK) { delete g; return st; }
  if (num_nodes > 0 && vmax >= num_nodes) {
    delete g;
    return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_create: node id >= num_nodes");
  }
  g->device = hm::current_device();
  g->csr.E = num_edges;
  g->csr.V = num_nodes > 0 ? num_nodes : vmax + 1;
  auto cleanup = [&](hm_status s) { hm_graph_destroy(g); return s; };
  if (hipSetDevice(g->device) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(hm::fail(HM_ERR_DEVICE, "hipStreamCreate failed"));
  for (auto& e : g->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hipEventCreate failed"));
  st = hm::upload_csr(g, h_off, h_nodes);
  g->live.assign((size_t)num_edges, 1);
  if (st != HM_OK) return cleanup(st);
  *out = g;
  return HM_OK;
}

hm_status hm_graph_destroy(hm_graph* g) {
  if (!g) return HM_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  g->scratch.release();
  g->dyn.release();
  g->csr.release();
  for (auto& e : g->ev)
    if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return HM_OK;
}

hm_status hm_count_static(hm_graph* g, uint64_t out[HM_STATIC_BINS]) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_count_static: null argument");
  return hm::count_static(g, out);
}

hm_status hm_graph_get_stats(hm_graph* g, hm_graph_stats* out) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_stats: null argument");
  if (!g->have_stats) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_stats: no count has run");
  *out = g->stats;
  return HM_OK;
}

hm_status hm_graph_phase_ms(hm_graph* g, const char* phase, double* ms) {
  hm::clear_error();
  if (!g || !phase || !ms) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_phase_ms: null argument");
  if (std::strcmp(phase, "project") == 0) *ms = g->ms_project;
  else if (std::strcmp(phase, "triangles") == 0) *ms = g->ms_triangles;
  else return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_phase_ms: unknown phase");
  return HM_OK;
}

}  // extern "C"
