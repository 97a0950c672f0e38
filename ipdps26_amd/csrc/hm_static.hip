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
//                 M[c] != 0 <=> triangle, C_ac = popc(M[c]), g_abc = popc(M[b] & M[c]); one
//                 LDS atomic per triangle into the slot that fixes its closed bin and its three
//                 as-if-open bins, the slots folded into the bins once per centre.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_static_dense.hpp"
#include "hm_tables.hpp"

namespace hm {

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_count(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, uint32_t* __restrict__ deg, uint32_t* __restrict__ dlow,
    unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t a = blockIdx.x;
  proj_count_item<MT>(reinterpret_cast<MT*>(smem), E, a, eoff, enodes, esize, voff, vedges, deg + a,
                      dlow + a, acc);
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_fill(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const int32_t* __restrict__ rank, const unsigned long long* __restrict__ roff,
    uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  proj_fill_item<MT>(reinterpret_cast<MT*>(smem), E, (int32_t)blockIdx.x, eoff, enodes, voff, vedges,
                     rank, roff, rlen, rows);
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

// Forms of the one-table triangle kernel: waves per centre workgroup x copies of the 243 uint32
// slot counters, shared by the workgroup (copy = lane % copies). Measured on the 5,000-hyperedge
// graph: 8 waves 4.45 / 4.46 ms with 8 / 16 copies, 4 waves 5.41 / 5.38 ms. Small graphs prefer 4
// waves; the form is autotuned per shape (below). The tiled kernel keeps 8 waves, 16 copies.
constexpr int kTriWaves = 8;
struct TriFormSpec {
  int waves, copies;
};
constexpr TriFormSpec kTriFormSpecs[] = {{4, 8}, {4, 16}, {8, 8}, {8, 16}};
constexpr int kNumTriFormSpecs = 4;
constexpr int kDefaultTriForm = 3;  // 8 waves, 16 copies

template <typename MT, int NW, int NC>
__global__ __launch_bounds__(NW * kWave) void k_tri(
    int32_t E, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vedges,
    const int32_t* __restrict__ rank, const int32_t* __restrict__ perm,
    const uint8_t* __restrict__ szr, const unsigned long long* __restrict__ roff,
    const uint32_t* __restrict__ rlen, const uint32_t* __restrict__ rows,
    const int32_t* __restrict__ order, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  tri_item<MT, NW, NC>(reinterpret_cast<MT*>(smem), E, order[blockIdx.x], eoff, enodes, voff, vedges,
                       rank, perm, szr, roff, rlen, rows, acc);
}

// ---------------------------------------------------------------------------------------------
// Tiled kernels: graphs whose hyperedge ids do not fit one LDS table. The key space (ids, or
// degree ranks) is cut into tiles of T keys; a workgroup walks the tiles of its hyperedge a,
// building the mask table of one tile at a time from the part of each node's V2H row that falls
// in the tile (binary search, rows ascending) and visiting only the entries it set: an entry x is
// visited by the lane holding the lowest node of a in x, which also clears it. The cost per
// (a, tile) is the tile's incidences plus |a| binary searches, never the tile size.
// ---------------------------------------------------------------------------------------------

// Projected degree, lower degree and neighbour classes of a, over id tiles.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_count_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint8_t* __restrict__ esize, const uint32_t* __restrict__ voff,
    const int32_t* __restrict__ vedges, uint32_t* __restrict__ deg, uint32_t* __restrict__ dlow,
    unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t cls[2 * HM_MAX_EDGE_SIZE];
  __shared__ uint32_t dred[2][kWavesPerBlock];
  __shared__ TileRanges tr;
  const int32_t a = blockIdx.x;
  const uint32_t s = eoff[a];
  const int sa = (int)(eoff[a + 1] - s);
  for (int k = threadIdx.x; k < 2 * HM_MAX_EDGE_SIZE; k += blockDim.x) cls[k] = 0;
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  uint32_t d = 0, dl = 0;
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
        const int C = popc(m);
        atomicAdd(&cls[2 * (C - 1) + (esize[x] > C ? 1 : 0)], 1u);
        ++d;
        dl += x < a ? 1u : 0u;
        M[x - t0] = (MT)0;
      }
    });
  }
  d = wave_sum(d);
  dl = wave_sum(dl);
  if (lane_id() == 0) {
    dred[0][wave_id()] = d;
    dred[1][wave_id()] = dl;
  }
  __syncthreads();
  open_all_pairs(cls, sa, acc);
  if (threadIdx.x == 0) {
    uint32_t dd = 0, ll = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) {
      dd += dred[0][w];
      ll += dred[1][w];
    }
    deg[a] = dd;
    dlow[a] = ll;
  }
}

// Upper row of a over rank tiles (vrank = V2H relabelled to ranks, ascending per node): entries
// (rank << 7 | C) grouped by tile in ascending tile order, in any order within a tile.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_proj_fill_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vrank,
    const int32_t* __restrict__ rank, const unsigned long long* __restrict__ roff,
    uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ TileRanges tr;
  __shared__ uint32_t s_cnt;
  const int32_t a = blockIdx.x;
  const int32_t ra = rank[a];
  const uint32_t s = eoff[a];
  const int sa = (int)(eoff[a + 1] - s);
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;
  if (threadIdx.x == 0) s_cnt = 0;
  uint32_t* out = rows + roff[ra];
  for (int32_t t0 = ((ra + 1) / T) * T; t0 < E; t0 += T) {
    const int32_t t1 = min(E, t0 + T);
    const uint32_t total = tile_ranges(tr, sa, enodes + s, voff, vrank, max(t0, ra + 1), t1);
    if (total == 0) continue;
    for_tile_entries(tr, sa, total, vrank,
                     [&](int i, int32_t x) { atomicOr(&M[x - t0], (MT)1 << i); });
    __syncthreads();
    for_tile_entries(tr, sa, total, vrank, [&](int i, int32_t x) {
      const MT m = M[x - t0];
      if (m && low_bit(m) == i) {
        out[atomicAdd(&s_cnt, 1u)] = ((uint32_t)x << kCntBits) | (uint32_t)popc(m);
        M[x - t0] = (MT)0;
      }
    });
  }
  __syncthreads();
  if (threadIdx.x == 0) rlen[ra] = s_cnt;
}

// Triangles of centre rank r over rank tiles of c. For each b in r's upper row, state at
// uoff[r] + ib keeps a's mask in b (computed once) and the read position in b's row, which is
// grouped by tile, so the entries of b's row in the current tile are the next ones from there.
// Tiles without work are skipped: the next tile is the smallest one any cursor points into.
template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_tri_tiled(
    int32_t E, int32_t T, const uint32_t* __restrict__ eoff, const int32_t* __restrict__ enodes,
    const uint32_t* __restrict__ voff, const int32_t* __restrict__ vrank,
    const int32_t* __restrict__ perm, const uint8_t* __restrict__ szr,
    const unsigned long long* __restrict__ roff, const uint32_t* __restrict__ rlen,
    const uint32_t* __restrict__ rows, const unsigned long long* __restrict__ uoff,
    uint32_t* __restrict__ cur, MT* __restrict__ pm, const int32_t* __restrict__ order,
    unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  uint8_t* szt = reinterpret_cast<uint8_t*>(M + T);
  constexpr int kCopies = 16;  // slot counter copies shared by the workgroup (copy = lane % 16)
  __shared__ uint32_t hs[kTriSlots * kCopies];
  __shared__ unsigned long long hb[32];
  __shared__ int32_t s_an[HM_MAX_EDGE_SIZE];
  __shared__ TileRanges tr;
  __shared__ uint32_t s_end;
  __shared__ int32_t s_next;
  const int32_t r = order[blockIdx.x];
  const uint32_t nA = rlen[r];
  if (nA < 2) return;
  const int32_t a = perm[r];
  const uint32_t s = eoff[a];
  const int sa = (int)(eoff[a + 1] - s);
  for (int k = threadIdx.x; k < kTriSlots * kCopies; k += blockDim.x) hs[k] = 0u;
  if (threadIdx.x < 32) hb[threadIdx.x] = 0ull;
  if ((int)threadIdx.x < sa) s_an[threadIdx.x] = enodes[s + threadIdx.x];
  for (int32_t x = threadIdx.x; x < T; x += blockDim.x) M[x] = (MT)0;

  const int lane = lane_id(), w = wave_id();
  uint32_t* hc = hs + (lane & (kCopies - 1));
  const uint32_t* rowA = rows + roff[r];
  const unsigned long long st = uoff[r];
  uint32_t prev_end = 0;
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
    const uint32_t total = tile_ranges(tr, sa, enodes + s, voff, vrank, max(t0, r + 1), t1);
    if (threadIdx.x == 0) s_next = INT32_MAX;
    for_tile_entries(tr, sa, total, vrank, [&](int i, int32_t x) {
      atomicOr(&M[x - t0], (MT)1 << i);
      szt[x - t0] = szr[x];
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
          mab = pair_mask<MT>(s_an, sa, eoff, enodes, perm[rb]);
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
        const int sbL = szr[rbL];
        const uint32_t* rowB = rows + obL;
        // kTriLoads chunks of b's row per step; the row is ascending, so the entries inside the
        // tile are a prefix and each chunk's valid entries follow the previous chunk's
        for (;;) {
          uint32_t e[kTriLoads];
          bool v[kTriLoads];
#pragma unroll
          for (int u = 0; u < kTriLoads; ++u) {
            const uint32_t j = pos + (uint32_t)(u * kWave) + lane;
            e[u] = j < nBL ? rowB[j] : 0xFFFFFFFFu;
            v[u] = j < nBL && (int32_t)(e[u] >> kCntBits) < t1;
          }
          uint32_t nv = 0;
#pragma unroll
          for (int u = 0; u < kTriLoads; ++u) {
            if (v[u]) {
              const int32_t rc = (int32_t)(e[u] >> kCntBits);
              const MT mc = M[rc - t0];
              if (mc) {
                const int cbc = (int)(e[u] & kCntMask);
                const int cac = popc(mc);
                const int g = popc((MT)(mabL & mc));
                const int sc = szt[rc - t0];
                atomicAdd(&hc[tri_slot(cabL - sa, cabL - sbL, cabL, sc, cbc, cac, g) * kCopies], 1u);
              }
            }
            nv += (uint32_t)__popcll(__ballot(v[u]));
          }
          pos += nv;
          if (nv < (uint32_t)(kTriLoads * kWave)) break;
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
    // fold this tile's slots into the bins and flush them (bounds the uint32 counters)
    for (int p = threadIdx.x; p < kTriSlots; p += blockDim.x) {
      unsigned long long n = 0;
      for (int c = 0; c < kCopies; ++c) {
        n += hs[p * kCopies + c];
        hs[p * kCopies + c] = 0u;
      }
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
    if (threadIdx.x < kStaticBins && hb[threadIdx.x]) {
      atomicAdd(&acc[kAccHist + threadIdx.x], hb[threadIdx.x]);
      hb[threadIdx.x] = 0ull;
    }
    for_tile_entries(tr, sa, total, vrank, [&](int, int32_t x) { M[x - t0] = (MT)0; });
    const int32_t kn = s_next;
    prev_end = end;
    if (kn == INT32_MAX) break;
    k = kn;
  }
}

// Node id of every V2H position: node[k] = v for voff[v] <= k < voff[v + 1].
__global__ void k_inc_node_keys(int32_t V, int64_t I, const uint32_t* __restrict__ voff,
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

__global__ void k_low_words(const unsigned long long* __restrict__ keys, int32_t* __restrict__ out,
                            int64_t I) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < I) out[k] = (int32_t)(uint32_t)keys[k];
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
                                                  const uint32_t* __restrict__ dlow,
                                                  const uint32_t* __restrict__ rlen,
                                                  unsigned long long* __restrict__ acc) {
  unsigned long long p = 0, mx = 0, wd = 0, up = 0, orr = 0;
  for (int32_t a = blockIdx.x * blockDim.x + threadIdx.x; a < E; a += gridDim.x * blockDim.x) {
    const unsigned long long d = deg[a];
    p += d;
    mx = d > mx ? d : mx;
    wd += d * (d - (d > 0)) / 2;
    up += rlen[a];
    orr += d * dlow[a];  // sum_a sum_{b in N(a), b > a} d_b = sum_b d_b #{a in N(b), a < b}
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

// LDS reserved for the static arrays of the triangle kernels (16 KB histogram + tables), next to
// the dynamic mask table.
constexpr size_t kTiledStaticLds = 20 * 1024;

// Largest tile (multiple of 64 ids) whose table and size bytes fit the LDS limit.
template <typename MT>
static int32_t auto_tile(int lim) {
  const size_t avail = (size_t)lim > kTiledStaticLds ? (size_t)lim - kTiledStaticLds : 0;
  return (int32_t)((avail / (sizeof(MT) + 1)) & ~(size_t)63);
}

// Autotuned one-table triangle forms, per process: (device, mask bytes, log2 E) -> form index.
static std::mutex g_tri_mu;
static std::map<std::tuple<int, int, int>, int> g_tri_tuned;

static int ceil_log2_i(int64_t x) {
  int k = 0;
  while (((int64_t)1 << k) < x) ++k;
  return k;
}

template <typename MT, int NW, int NC>
static hm_status launch_tri_one(hm_graph* g, size_t lds, unsigned long long* acc) {
  HM_TRY(set_lds_attr<MT>((const void*)k_tri<MT, NW, NC>, lds));
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  k_tri<MT, NW, NC><<<c.E, NW * kWave, lds, g->stream>>>(
      c.E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
      s.rank.as<int32_t>(), s.perm.as<int32_t>(), s.szr.as<uint8_t>(), s.roff.as<unsigned long long>(),
      s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(), s.order.as<int32_t>(), acc);
  HM_HIP(hipGetLastError());
  return HM_OK;
}

template <typename MT>
static hm_status launch_tri_form(hm_graph* g, int form, size_t lds, unsigned long long* acc) {
  switch (form) {
    case 0: return launch_tri_one<MT, 4, 8>(g, lds, acc);
    case 1: return launch_tri_one<MT, 4, 16>(g, lds, acc);
    case 2: return launch_tri_one<MT, 8, 8>(g, lds, acc);
    default: return launch_tri_one<MT, 8, 16>(g, lds, acc);
  }
}

// Form of the one-table triangle kernel for g: forced, or tuned for g's shape, or tuned now (every
// form once into scratch, the bins checked equal, the fastest kept for the process). capture:
// the form of the eager count just made.
template <typename MT>
static hm_status pick_tri_form(hm_graph* g, size_t lds, bool capture, int* form_out) {
  if (capture) {
    *form_out = g->last_tri_form >= 0 ? g->last_tri_form : kDefaultTriForm;
    return HM_OK;
  }
  g->last_tri_tuned = false;
  int form = g->tri_form;
  const auto key = std::make_tuple(g->device, (int)sizeof(MT), ceil_log2_i(g->csr.E));
  if (form < 0) {
    std::lock_guard<std::mutex> lk(g_tri_mu);
    if (g->tri_retune) g_tri_tuned.erase(key);
    auto it = g_tri_tuned.find(key);
    if (it != g_tri_tuned.end()) form = it->second;
  }
  g->tri_retune = false;
  if (form < 0 && alloc_forbidden()) form = kDefaultTriForm;  // tuning needs its own buffer
  if (form < 0) {
    CountScratch& s = g->scratch;
    HM_TRY(s.acc2.ensure(sizeof(unsigned long long) * kAccCount));
    unsigned long long* acc2 = s.acc2.as<unsigned long long>();
    unsigned long long first[kStaticBins], cur[kStaticBins];
    float best = 0.f;
    for (int f = 0; f < kNumTriFormSpecs; ++f) {
      k_zero_u64<<<1, 64, 0, g->stream>>>(acc2, kAccCount);
      HM_HIP(hipEventRecord(g->ev[4], g->stream));
      HM_TRY(launch_tri_form<MT>(g, f, lds, acc2));
      HM_HIP(hipEventRecord(g->ev[5], g->stream));
      HM_HIP(hipMemcpyAsync(cur, acc2, sizeof(cur), hipMemcpyDeviceToHost, g->stream));
      HM_HIP(hipStreamSynchronize(g->stream));
      float ms = 0.f;
      HM_HIP(hipEventElapsedTime(&ms, g->ev[4], g->ev[5]));
      if (f == 0) {
        std::memcpy(first, cur, sizeof(cur));
      } else if (std::memcmp(first, cur, sizeof(cur)) != 0) {
        set_error("static count: triangle forms %d and 0 disagree", f);
        return HM_ERR_DEVICE;
      }
      if (f == 0 || ms < best) {
        best = ms;
        form = f;
      }
    }
    std::lock_guard<std::mutex> lk(g_tri_mu);
    g_tri_tuned[key] = form;
    g->last_tri_tuned = true;
  }
  *form_out = form;
  return HM_OK;
}

// Static count. T == 0: one LDS table over all ids (dense kernels); T > 0: tiles of T ids/ranks.
// capture: only enqueue the launches (the stream is being captured, the graph is unchanged since
// the eager run that sized every buffer): no host waits, no events, no result copy.
template <typename MT>
static hm_status count_static_impl(hm_graph* g, unsigned long long* h_acc, int32_t T, bool capture) {
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  const int32_t E = c.E;
  const int64_t I = c.I;
  hipStream_t st = g->stream;
  const bool tiled = T > 0;
  const size_t lds_cnt = tiled ? dense_lds_bytes<MT>(T, false) : dense_lds_bytes<MT>(E, false);
  const size_t lds_tri = tiled ? dense_lds_bytes<MT>(T, true) : dense_lds_bytes<MT>(E, true);
  if (tiled) {
    HM_TRY(set_lds_attr<MT>((const void*)k_proj_count_tiled<MT>, lds_cnt));
    HM_TRY(set_lds_attr<MT>((const void*)k_proj_fill_tiled<MT>, lds_cnt));
    HM_TRY(set_lds_attr<MT>((const void*)k_tri_tiled<MT, kTriWaves>, lds_tri));
  } else {
    HM_TRY(set_lds_attr<MT>((const void*)k_proj_count<MT>, lds_cnt));
    HM_TRY(set_lds_attr<MT>((const void*)k_proj_fill<MT>, lds_cnt));
  }

  HM_TRY(s.deg.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.dlow.ensure(sizeof(uint32_t) * E));
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
  HM_TRY(s.acc.ensure(sizeof(unsigned long long) * kAccCount));
  if (tiled && I >= ((int64_t)1 << 31))
    return fail(HM_ERR_OVER_BUDGET, "static count: tiled path supports fewer than 2^31 incidences");
  if (tiled) {
    HM_TRY(s.vkeys.ensure(sizeof(unsigned long long) * std::max<int64_t>(I, 1)));
    HM_TRY(s.vkeys2.ensure(sizeof(unsigned long long) * std::max<int64_t>(I, 1)));
    HM_TRY(s.vrank.ensure(sizeof(int32_t) * std::max<int64_t>(I, 1)));
    HM_TRY(s.uoff.ensure(sizeof(unsigned long long) * (E + 1)));
  }

  unsigned long long* acc = s.acc.as<unsigned long long>();
  const unsigned eb = (unsigned)((E + kBlock - 1) / kBlock);

  if (!capture) HM_HIP(hipEventRecord(g->ev[0], st));
  k_zero_u64<<<1, 64, 0, st>>>(acc, kAccCount);
  // 1. projected degrees + combinatorial open counts
  if (tiled)
    k_proj_count_tiled<MT><<<E, kBlock, lds_cnt, st>>>(
        E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
        c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), s.deg.as<uint32_t>(), s.dlow.as<uint32_t>(), acc);
  else
    k_proj_count<MT><<<E, kBlock, lds_cnt, st>>>(
        E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(), c.voff.as<uint32_t>(),
        c.vedges.as<int32_t>(), s.deg.as<uint32_t>(), s.dlow.as<uint32_t>(), acc);
  HM_HIP(hipGetLastError());
  // 2. degree order: stable radix sort of (d_a, a)
  k_iota<<<eb, kBlock, 0, st>>>(s.ids.as<int32_t>(), E);
  int vbits = 1;
  while (vbits < 31 && ((int64_t)1 << vbits) <= c.V) ++vbits;
  size_t tb1 = 0, tb2 = 0, tb3 = 0, tb4 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, s.deg.as<uint32_t>(), s.work2.as<uint32_t>(),
                                            s.ids.as<int32_t>(), s.perm.as<int32_t>(), E, 0, 32, st));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, s.degr.as<unsigned long long>(),
                                          s.roff.as<unsigned long long>(), E + 1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb3, s.work.as<unsigned long long>(),
                                                      s.work2.as<unsigned long long>(),
                                                      s.ranks.as<int32_t>(), s.order.as<int32_t>(),
                                                      E, 0, 64, st));
  if (tiled)
    HM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb4, s.vkeys.as<unsigned long long>(),
                                             s.vkeys2.as<unsigned long long>(), (int)I, 0, 32 + vbits, st));
  HM_TRY(s.tmp.ensure(std::max(std::max(tb1, tb2), std::max(tb3, tb4))));
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
  if (tiled && I > 0) {  // V2H relabelled to ranks, ascending within every node's row
    const unsigned ib = (unsigned)((I + kBlock - 1) / kBlock);
    k_inc_node_keys<<<ib, kBlock, 0, st>>>(c.V, I, c.voff.as<uint32_t>(), c.vedges.as<int32_t>(),
                                           s.rank.as<int32_t>(), s.vkeys.as<unsigned long long>());
    tb = s.tmp.bytes;
    HM_HIP(hipcub::DeviceRadixSort::SortKeys(s.tmp.p, tb, s.vkeys.as<unsigned long long>(),
                                             s.vkeys2.as<unsigned long long>(), (int)I, 0, 32 + vbits, st));
    k_low_words<<<ib, kBlock, 0, st>>>(s.vkeys2.as<unsigned long long>(), s.vrank.as<int32_t>(), I);
  }
  if (!capture) {
    unsigned long long total_cap = 0, dmax = 0;
    HM_HIP(hipMemcpyAsync(&total_cap, s.roff.as<unsigned long long>() + E, sizeof(total_cap),
                          hipMemcpyDeviceToHost, st));
    // degr holds the degrees in ascending order (the sorted keys): its last entry is d_max
    HM_HIP(hipMemcpyAsync(&dmax, s.degr.as<unsigned long long>() + (E - 1), sizeof(dmax),
                          hipMemcpyDeviceToHost, st));
    HM_HIP(hipStreamSynchronize(st));
    // k_tri_tiled's uint32 slot counters are flushed per tile; copy (lane % 16) of a centre's
    // workgroup takes at most a 16th of every neighbour row's entries inside the tile, plus one
    // partial step per neighbour: at most d_max * (T / 16 + 4 * kTriLoads + 1) per tile
    if (tiled && dmax * ((unsigned long long)T / 16 + 4 * kTriLoads + 1) >= (1ull << 32)) {
      set_error("static count: projected degree %llu x tile of %d ids exceeds the bound of the "
                "32-bit slot counters (2^32)", dmax, T);
      return HM_ERR_UNSUPPORTED;
    }
    HM_TRY(s.rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total_cap, 1)));
  }
  // 4. upper rows
  if (tiled)
    k_proj_fill_tiled<MT><<<E, kBlock, lds_cnt, st>>>(
        E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(),
        s.vrank.as<int32_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
        s.rlen.as<uint32_t>(), s.rows.as<uint32_t>());
  else
    k_proj_fill<MT><<<E, kBlock, lds_cnt, st>>>(
        E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(),
        c.vedges.as<int32_t>(), s.rank.as<int32_t>(), s.roff.as<unsigned long long>(),
        s.rlen.as<uint32_t>(), s.rows.as<uint32_t>());
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
  if (tiled) {  // per (a, b) state of the tiled triangle kernel, at the scan of the upper lengths
    k_widen<<<(unsigned)((E + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
        s.rlen.as<uint32_t>(), s.degr.as<unsigned long long>(), E);
    tb = s.tmp.bytes;
    HM_HIP(hipcub::DeviceScan::ExclusiveSum(s.tmp.p, tb, s.degr.as<unsigned long long>(),
                                            s.uoff.as<unsigned long long>(), E + 1, st));
    if (!capture) {
      unsigned long long U = 0;
      HM_HIP(hipMemcpyAsync(&U, s.uoff.as<unsigned long long>() + E, sizeof(U), hipMemcpyDeviceToHost, st));
      HM_HIP(hipStreamSynchronize(st));
      HM_TRY(s.cur.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(U, 1)));
      HM_TRY(s.pm.ensure(sizeof(MT) * (size_t)std::max<unsigned long long>(U, 1)));
    }
  }
  int form = -1;
  if (!tiled) HM_TRY(pick_tri_form<MT>(g, lds_tri, capture, &form));
  g->last_tri_form = form;
  if (!capture) HM_HIP(hipEventRecord(g->ev[1], st));
  // 6. triangles
  if (tiled)
    k_tri_tiled<MT, kTriWaves><<<E, kTriWaves * kWave, lds_tri, st>>>(
        E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.voff.as<uint32_t>(),
        s.vrank.as<int32_t>(), s.perm.as<int32_t>(), s.szr.as<uint8_t>(),
        s.roff.as<unsigned long long>(), s.rlen.as<uint32_t>(), s.rows.as<uint32_t>(),
        s.uoff.as<unsigned long long>(), s.cur.as<uint32_t>(), s.pm.as<MT>(), s.order.as<int32_t>(), acc);
  else
    HM_TRY(launch_tri_form<MT>(g, form, lds_tri, acc));
  HM_HIP(hipGetLastError());
  if (!capture) HM_HIP(hipEventRecord(g->ev[2], st));
  k_stats<<<64, kBlock, 0, st>>>(E, c.V, c.voff.as<uint32_t>(), s.deg.as<uint32_t>(),
                                 s.dlow.as<uint32_t>(), s.rlen.as<uint32_t>(), acc);
  g->last_tile = T;
  if (capture) return HM_OK;
  HM_HIP(hipMemcpyAsync(h_acc, acc, sizeof(unsigned long long) * kAccCount, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_HIP(hipEventElapsedTime(&g->ms_project, g->ev[0], g->ev[1]));
  HM_HIP(hipEventElapsedTime(&g->ms_triangles, g->ev[1], g->ev[2]));
  return HM_OK;
}

using CountImpl = hm_status (*)(hm_graph*, unsigned long long*, int32_t, bool);

// Replays the captured count of g: one graph launch, one wait. The phases are not separated:
// ms_triangles receives the whole count, ms_project 0.
static hm_status replay_count(hm_graph* g, unsigned long long* h_acc) {
  hipStream_t st = g->stream;
  HM_HIP(hipEventRecord(g->ev[0], st));
  HM_HIP(hipGraphLaunch(g->replay.exec, st));
  HM_HIP(hipEventRecord(g->ev[2], st));
  HM_HIP(hipMemcpyAsync(h_acc, g->scratch.acc.p, sizeof(unsigned long long) * kAccCount,
                        hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  g->ms_project = 0.f;
  HM_HIP(hipEventElapsedTime(&g->ms_triangles, g->ev[0], g->ev[2]));
  return HM_OK;
}

// Captures the count of g (just counted eagerly, so every buffer is sized) for later replays.
static hm_status capture_count(hm_graph* g, CountImpl impl, int32_t T) {
  g->replay.reset();
  CaptureGuard capturing;  // no allocation or free on another thread meanwhile
  HM_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  const hm_status st = impl(g, nullptr, T, true);
  CapturedGraph cap;
  const hm_status st2 = end_capture(g->stream, &cap);  // ends the capture on every path
  if (st != HM_OK || st2 != HM_OK) {
    cap.reset();
    return st != HM_OK ? st : st2;
  }
  g->replay = cap;  // exec null when the capture holds a memset node: counts stay eager
  g->replay_version = g->version;
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

int64_t graph_device_bytes(const hm_graph* g) {
  int64_t n = 0;
  auto add = [&](const DevBuf& b) { n += b.p ? (int64_t)b.bytes : 0; };
  const DeviceCsr& c = g->csr;
  for (const DevBuf* b : {&c.eoff, &c.enodes, &c.esize, &c.voff, &c.vedges}) add(*b);
  const CountScratch& s = g->scratch;
  for (const DevBuf* b : {&s.deg, &s.ids, &s.perm, &s.rank, &s.degr, &s.roff, &s.rlen, &s.rows, &s.szr,
                          &s.work, &s.work2, &s.order, &s.ranks, &s.dlow, &s.acc, &s.acc2, &s.tmp,
                          &s.vkeys, &s.vkeys2, &s.vrank, &s.uoff, &s.cur, &s.pm})
    add(*b);
  const DynScratch& d = g->dyn;
  for (const DevBuf* b : {&d.deg, &d.roff, &d.rows, &d.flag, &d.ids, &d.acc, &d.tmp, &d.ins_off, &d.ins_nodes,
                          &d.sizes, &d.noff, &d.nnodes, &d.neoff, &d.nsize, &d.inc_node, &d.inc_edge,
                          &d.srt_node, &d.srt_edge, &d.nvoff, &d.slen, &d.soff, &d.scur, &d.spm, &d.ord})
    add(*b);
  add(g->uplan.in);
  g->hub.each([&](const DevBuf& b) { add(b); });
  return n;
}

// Projection statistics of g (the kAccP .. kAccMaxNodeDeg slots of h_acc) from the projected and
// lower degrees alone: k_proj_count(_tiled) and k_stats, no rows and no triangles (upper-row
// total 0). The hub-aware count reports them beside its bins.
hm_status projection_stats(hm_graph* g, unsigned long long* h_acc) {
  DeviceCsr& c = g->csr;
  CountScratch& s = g->scratch;
  const int32_t E = c.E;
  hipStream_t st = g->stream;
  const int lim = lds_limit_bytes(g->device);
  const bool wide = c.max_size > 32;
  HM_TRY(s.deg.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.dlow.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.rlen.ensure(sizeof(uint32_t) * E));
  HM_TRY(s.acc2.ensure(sizeof(unsigned long long) * kAccCount));
  unsigned long long* acc = s.acc2.as<unsigned long long>();
  k_zero_u64<<<1, 64, 0, st>>>(acc, kAccCount);
  HM_HIP(hipMemsetAsync(s.rlen.p, 0, sizeof(uint32_t) * E, st));
  auto run = [&](auto mt) -> hm_status {
    using MT = decltype(mt);
    const size_t dense = dense_lds_bytes<MT>(E, false);
    if (dense + kTiledStaticLds <= (size_t)lim) {
      HM_TRY(set_lds_attr<MT>((const void*)k_proj_count<MT>, dense));
      k_proj_count<MT><<<E, kBlock, dense, st>>>(E, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                                  c.esize.as<uint8_t>(), c.voff.as<uint32_t>(),
                                                  c.vedges.as<int32_t>(), s.deg.as<uint32_t>(),
                                                  s.dlow.as<uint32_t>(), acc);
    } else {
      const int32_t T = auto_tile<MT>(lim);
      const size_t lds = dense_lds_bytes<MT>(T, false);
      HM_TRY(set_lds_attr<MT>((const void*)k_proj_count_tiled<MT>, lds));
      k_proj_count_tiled<MT><<<E, kBlock, lds, st>>>(E, T, c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(),
                                                      c.esize.as<uint8_t>(), c.voff.as<uint32_t>(),
                                                      c.vedges.as<int32_t>(), s.deg.as<uint32_t>(),
                                                      s.dlow.as<uint32_t>(), acc);
    }
    HM_HIP(hipGetLastError());
    return HM_OK;
  };
  HM_TRY(wide ? run(mask64_t{}) : run(uint32_t{}));
  k_stats<<<64, kBlock, 0, st>>>(E, c.V, c.voff.as<uint32_t>(), s.deg.as<uint32_t>(), s.dlow.as<uint32_t>(),
                                 s.rlen.as<uint32_t>(), acc);
  HM_HIP(hipGetLastError());
  unsigned long long tmp[kAccCount];
  HM_HIP(hipMemcpyAsync(tmp, acc, sizeof(tmp), hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  for (int k : {kAccP, kAccMaxDeg, kAccOrient, kAccWedges, kAccUpper, kAccMaxNodeDeg}) h_acc[k] = tmp[k];
  return HM_OK;
}

// Graph statistics of the count just made (hm_graph_get_stats).
static void fill_count_stats(hm_graph* g, const uint64_t out[HM_STATIC_BINS], const unsigned long long* acc) {
  const int32_t E = g->csr.E;
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
  g->stats_lazy = false;
}

// ---------------------------------------------------------------------------------------------
// Static count without allocation (hm_count_static_noalloc): the one-table pipeline above with its
// scratch carved from `scratch`, one file-scope device array per GPU. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol): its device address
// comes from hipGetSymbolAddress, looked up once per device with that device current.
constexpr size_t kNoAllocScratchBytes = (size_t)HM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
  std::mutex mu;         // serializes the calls on one device (they share its scratch)
  void* ptr = nullptr;   // device address of that device's copy of `scratch`
};
constexpr int kNoAllocMaxDevices = 64;
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// Device address of `scratch` on device dev (current); dev's mutex is held.
static hm_status noalloc_base(int dev, unsigned char** base) {
  NoAllocDevice& d = g_noalloc[dev];
  if (!d.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!p) return fail(HM_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
    d.ptr = p;
  }
  *base = static_cast<unsigned char*>(d.ptr);
  return HM_OK;
}

// Carves the one-table count's buffers for E hyperedges out of [base, base + cap): the
// per-hyperedge arrays, the accumulator and the sort/scan storage first, the packed projection
// rows in the rest. Fails (nothing launched) when the fixed part does not fit.
static hm_status noalloc_carve(hm_graph* g, unsigned char* base, size_t cap, CountScratch* s) {
  const int32_t E = g->csr.E;
  size_t at = 0;
  auto take = [&](DevBuf& b, size_t bytes) {
    at = (at + 255) & ~(size_t)255;
    b.p = at + bytes <= cap ? base + at : nullptr;
    b.bytes = bytes;
    at += bytes;
  };
  const size_t e4 = sizeof(uint32_t) * (size_t)E, e8 = sizeof(unsigned long long) * (size_t)E;
  take(s->deg, e4);
  take(s->dlow, e4);
  take(s->ids, e4);
  take(s->perm, e4);
  take(s->rank, e4);
  take(s->degr, e8 + 8);
  take(s->roff, e8 + 8);
  take(s->rlen, e4);
  take(s->szr, (size_t)E);
  take(s->work, e8);
  take(s->work2, e8);
  take(s->order, e4);
  take(s->ranks, e4);
  take(s->acc, sizeof(unsigned long long) * kAccCount);
  // sort and scan storage, sized as the pipeline asks (queries only: nothing runs)
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                            (const int32_t*)nullptr, (int32_t*)nullptr, E, 0, 32, g->stream));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (const unsigned long long*)nullptr,
                                          (unsigned long long*)nullptr, E + 1, g->stream));
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb3, (const unsigned long long*)nullptr,
                                                      (unsigned long long*)nullptr, (const int32_t*)nullptr,
                                                      (int32_t*)nullptr, E, 0, 64, g->stream));
  take(s->tmp, std::max({tb1, tb2, tb3, (size_t)16}));
  at = (at + 255) & ~(size_t)255;
  if (at >= cap) {
    set_error("hm_count_static_noalloc: %d hyperedges need %zu bytes of scratch before the "
              "projection rows; the static scratch holds %zu", E, at, cap);
    return HM_ERR_OVER_BUDGET;
  }
  s->rows.p = base + at;  // the rest: 4 B per directed projection pair (checked before the fill)
  s->rows.bytes = cap - at;
  return HM_OK;
}

hm_status count_static_noalloc(hm_graph* g, uint64_t out[HM_STATIC_BINS]) {
  const int32_t E = g->csr.E;
  unsigned long long acc[kAccCount] = {0};
  g->ms_project = g->ms_triangles = 0.f;
  g->last_replayed = false;
  if (E >= 3) {
    if (g->csr.max_size > HM_MAX_EDGE_SIZE)
      return fail(HM_ERR_UNSUPPORTED, "hyperedge larger than 64 nodes");
    if (g->tile > 0)
      return fail(HM_ERR_UNSUPPORTED, "hm_count_static_noalloc: forced tiles need the allocating count");
    if (g->device < 0 || g->device >= kNoAllocMaxDevices)
      return fail(HM_ERR_UNSUPPORTED, "hm_count_static_noalloc: device index beyond the scratch table");
    HM_HIP(hipSetDevice(g->device));
    const int lim = lds_limit_bytes(g->device);
    const bool wide = g->csr.max_size > 32;
    CountImpl impl;
    if (!wide && dense_lds_bytes<uint32_t>(E, true) + kTiledStaticLds <= (size_t)lim)
      impl = count_static_impl<uint32_t>;
    else if (dense_lds_bytes<mask64_t>(E, true) + kTiledStaticLds <= (size_t)lim)
      impl = count_static_impl<mask64_t>;
    else
      return fail(HM_ERR_OVER_BUDGET, "hm_count_static_noalloc: the hyperedge ids exceed one LDS "
                                      "table (the tiled count allocates)");
    NoAllocDevice& dev = g_noalloc[g->device];
    std::lock_guard<std::mutex> lk(dev.mu);
    unsigned char* base = nullptr;
    HM_TRY(noalloc_base(g->device, &base));
    CountScratch carved;
    HM_TRY(noalloc_carve(g, base, kNoAllocScratchBytes, &carved));
    std::swap(g->scratch, carved);  // the handle's own scratch is put back below on every path
    hm_status st;
    {
      NoAllocScope no_alloc;
      st = impl(g, acc, 0, false);
    }
    // the scratch is shared by the device's next call: nothing of this one may still run
    const hipError_t drained = hipStreamSynchronize(g->stream);
    std::swap(g->scratch, carved);
    if (st != HM_OK) return st;
    HM_HIP(drained);
  }
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = acc[kAccHist + k];
  fill_count_stats(g, out, acc);
  return HM_OK;
}

hm_status count_static(hm_graph* g, uint64_t out[HM_STATIC_BINS]) {
  const int32_t E = g->csr.E;
  unsigned long long acc[kAccCount] = {0};
  g->ms_project = g->ms_triangles = 0.f;
  g->last_replayed = false;
  if (E >= 3) {
    if (g->csr.max_size > HM_MAX_EDGE_SIZE)
      return fail(HM_ERR_UNSUPPORTED, "hyperedge larger than 64 nodes");
    if ((int64_t)E >= kMaxIds) return fail(HM_ERR_UNSUPPORTED, "more than 2^22 - 1 hyperedges");
    HM_HIP(hipSetDevice(g->device));
    g->hub_stats.hubs = 0;
    if (g->hubs != 0) {
      // forced hub count: its result or its error; automatic: the cost model's choice, lowered by
      // a quarter while some hyperedge's neighbour keys overflow the largest global histogram
      // (917,504 keys), else the triangle path
      int32_t h = g->hubs;
      if (h < 0) {  // the cost model's choice is kept while the store is unchanged
        if (g->hub_auto_version != g->version || g->hub_auto_h <= 0) {
          HM_TRY(hub_choose(g, &h, nullptr));
          g->hub_auto_h = h;
          g->hub_auto_version = g->version;
        }
        h = g->hub_auto_h;
      }
      for (;;) {
        bool overflow = false;
        const hm_status st = count_static_hub(g, h, acc, &overflow, /*stats*/ false);
        if (st == HM_OK) {
          for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = g->lazy_bins[k] = acc[kAccHist + k];
          g->have_stats = true;  // completed by the first hm_graph_get_stats
          g->stats_lazy = true;
          return HM_OK;
        }
        if (g->hubs > 0 || !overflow) return st;
        if (h == 1) break;
        h = std::max<int32_t>(1, h * 3 / 4);
        g->hub_auto_h = h;
        clear_error();
      }
      g->hub_stats.hubs = 0;
      g->hub_auto_h = 0;
      for (int k = 0; k < kAccCount; ++k) acc[k] = 0;
      clear_error();
    }
    const int lim = lds_limit_bytes(g->device);
    const bool wide = g->csr.max_size > 32;
    CountImpl impl;
    int32_t T = 0;
    if (g->tile > 0) {
      const size_t need = (size_t)g->tile * ((wide ? 8 : 4) + 1) + kTiledStaticLds;
      if (need > (size_t)lim) {
        set_error("static count: a tile of %d ids exceeds the LDS limit (%d bytes)", g->tile, lim);
        return HM_ERR_BAD_INPUT;
      }
      impl = wide ? count_static_impl<mask64_t> : count_static_impl<uint32_t>;
      T = g->tile;
    } else if (!wide && dense_lds_bytes<uint32_t>(E, true) + kTiledStaticLds <= (size_t)lim) {
      impl = count_static_impl<uint32_t>;
    } else if (dense_lds_bytes<mask64_t>(E, true) + kTiledStaticLds <= (size_t)lim) {
      impl = count_static_impl<mask64_t>;
    } else {
      impl = wide ? count_static_impl<mask64_t> : count_static_impl<uint32_t>;
      T = wide ? auto_tile<mask64_t>(lim) : auto_tile<uint32_t>(lim);
    }
    if (g->replay_on && g->replay.exec && g->replay_version == g->version) {
      HM_TRY(replay_count(g, acc));
      g->last_replayed = true;
    } else {
      HM_TRY(impl(g, acc, T, false));
      // the counts are in hand: a capture that fails only leaves later counts eager
      if (g->replay_on && capture_count(g, impl, T) != HM_OK) clear_error();
    }
  }
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = acc[kAccHist + k];
  fill_count_stats(g, out, acc);
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
  g->replay.reset();
  g->uplan.release();
  g->scratch.release();
  g->dyn.release();
  g->hub.release();
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

hm_status hm_count_static_noalloc(hm_graph* g, uint64_t out[HM_STATIC_BINS]) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_count_static_noalloc: null argument");
  return hm::count_static_noalloc(g, out);
}

hm_status hm_graph_set_tile(hm_graph* g, int32_t tile_ids) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_set_tile: null handle");
  if (tile_ids < 0) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_set_tile: negative tile");
  if (g->tile != tile_ids) ++g->version;
  g->tile = tile_ids;
  return HM_OK;
}

hm_status hm_graph_set_hubs(hm_graph* g, int32_t hubs) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "null graph");
  if (hubs < -1 || hubs > hm::kHubMax) return hm::fail(HM_ERR_BAD_INPUT, "hubs must be -1, 0 or 1..23");
  g->hubs = hubs;
  return HM_OK;
}

hm_status hm_graph_get_hub_stats(hm_graph* g, hm_hub_stats* out) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "null argument");
  *out = g->hub_stats;
  return HM_OK;
}

hm_status hm_graph_count_hub_part(hm_graph* g, int32_t hubs, int64_t centre_begin, int64_t centre_end,
                                  int32_t other_parts, double seconds, uint64_t out[HM_STATIC_BINS],
                                  int64_t* light_triangles, int64_t* centre_done) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_count_hub_part: null argument");
  if (hubs < 1 || hubs > hm::kHubMax || (other_parts != 0 && other_parts != 1) || centre_begin < 0 ||
      (centre_end >= 0 && centre_end < centre_begin) || !(seconds >= 0))
    return hm::fail(HM_ERR_BAD_INPUT,
                    "hm_graph_count_hub_part: need hubs in 1..23, other_parts 0 or 1, 0 <= centre_begin <= "
                    "centre_end (or centre_end = -1) and seconds >= 0");
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = 0;
  if (light_triangles) *light_triangles = 0;
  const int32_t E = g->csr.E;
  if (centre_done) *centre_done = centre_end < 0 ? E : std::min<int64_t>(centre_end, E);
  g->hub_stats = hm_hub_stats{};
  if (E < 3) return HM_OK;
  if (g->csr.max_size > HM_MAX_EDGE_SIZE) return hm::fail(HM_ERR_UNSUPPORTED, "hyperedge larger than 64 nodes");
  if ((int64_t)E >= hm::kMaxIds) return hm::fail(HM_ERR_UNSUPPORTED, "more than 2^22 - 1 hyperedges");
  HM_HIP(hipSetDevice(g->device));
  hm::HubRun run;
  run.hubs_out = &g->hub_nodes;
  run.abc = other_parts != 0;
  run.stats = other_parts != 0;
  run.d_begin = centre_begin;
  run.d_end = centre_end;
  run.seconds = seconds;
  int64_t done = 0;
  run.d_done = &done;
  unsigned long long abc[hm::kAccCount] = {0}, d[hm::kAccCount] = {0};
  bool overflow = false;
  HM_TRY(hm::hub_parts(g, hubs, run, abc, d, &overflow));
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = abc[hm::kAccHist + k] + d[hm::kAccHist + k];
  if (light_triangles) *light_triangles = (int64_t)d[hm::kAccLightTri];
  if (centre_done) *centre_done = done;
  if (other_parts) {  // the projection statistics; the closed-triple total is not known from one part
    hm::fill_count_stats(g, out, abc);
    g->stats.triangles = -1;
  }
  return HM_OK;
}

hm_status hm_graph_set_tri_form(hm_graph* g, int32_t waves, int32_t copies) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_set_tri_form: null handle");
  int form = -2;
  if (waves == 0 || waves == -1) {
    form = -1;
  } else {
    for (int f = 0; f < hm::kNumTriFormSpecs; ++f)
      if (hm::kTriFormSpecs[f].waves == waves && hm::kTriFormSpecs[f].copies == copies) form = f;
  }
  if (form == -2)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_set_tri_form: waves must be 0, -1, 4 or 8 and copies 8 or 16");
  g->tri_form = form;
  g->tri_retune = waves == -1;
  ++g->version;  // a captured count holds the previous form
  return HM_OK;
}

hm_status hm_graph_get_tri_form(hm_graph* g, int32_t* waves, int32_t* copies, int32_t* tuned) {
  hm::clear_error();
  if (!g || !waves || !copies || !tuned) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_tri_form: null argument");
  const int f = g->last_tri_form;
  *waves = f >= 0 ? hm::kTriFormSpecs[f].waves : 0;
  *copies = f >= 0 ? hm::kTriFormSpecs[f].copies : 0;
  *tuned = g->last_tri_tuned ? 1 : 0;
  return HM_OK;
}

hm_status hm_graph_set_replay(hm_graph* g, int32_t on) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_set_replay: null handle");
  g->replay_on = on != 0;
  if (!g->replay_on) g->replay.reset();
  return HM_OK;
}

hm_status hm_graph_replay_state(hm_graph* g, int32_t* captured, int32_t* last_replayed) {
  hm::clear_error();
  if (!g || !captured || !last_replayed) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_replay_state: null argument");
  *captured = (g->replay.exec && g->replay_version == g->version) ? 1 : 0;
  *last_replayed = g->last_replayed ? 1 : 0;
  return HM_OK;
}

hm_status hm_graph_get_stats(hm_graph* g, hm_graph_stats* out) {
  hm::clear_error();
  if (!g || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_stats: null argument");
  if (!g->have_stats) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_stats: no count has run");
  if (g->stats_lazy) {  // the projection statistics of the last (hub-aware) count, from the degrees
    unsigned long long acc[hm::kAccCount] = {0};
    HM_HIP(hipSetDevice(g->device));
    HM_TRY(hm::projection_stats(g, acc));
    hm::fill_count_stats(g, g->lazy_bins, acc);
  }
  *out = g->stats;
  return HM_OK;
}

hm_status hm_graph_get_sizes(hm_graph* g, int32_t* num_edges, int32_t* num_nodes, int64_t* num_incidences) {
  hm::clear_error();
  if (!g || !num_edges || !num_nodes || !num_incidences)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_get_sizes: null argument");
  *num_edges = g->csr.E;
  *num_nodes = g->csr.V;
  *num_incidences = g->csr.I;
  return HM_OK;
}

hm_status hm_graph_read_csr(hm_graph* g, int64_t* offsets, int32_t* nodes, int64_t* voffsets,
                            int32_t* vedges) {
  hm::clear_error();
  if (!g) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_read_csr: null handle");
  const hm::DeviceCsr& c = g->csr;
  HM_HIP(hipSetDevice(g->device));
  HM_HIP(hipStreamSynchronize(g->stream));
  auto read_offsets = [&](const hm::DevBuf& src, int64_t n, int64_t* dst) -> hm_status {
    std::vector<uint32_t> tmp((size_t)n + 1);
    if (src.p) HM_HIP(hipMemcpy(tmp.data(), src.p, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k <= n; ++k) dst[k] = src.p ? (int64_t)tmp[(size_t)k] : 0;
    return HM_OK;
  };
  if (offsets) HM_TRY(read_offsets(c.eoff, c.E, offsets));
  if (voffsets) HM_TRY(read_offsets(c.voff, c.V, voffsets));
  if (nodes && c.I > 0) HM_HIP(hipMemcpy(nodes, c.enodes.p, sizeof(int32_t) * (size_t)c.I, hipMemcpyDeviceToHost));
  if (vedges && c.I > 0) HM_HIP(hipMemcpy(vedges, c.vedges.p, sizeof(int32_t) * (size_t)c.I, hipMemcpyDeviceToHost));
  return HM_OK;
}

hm_status hm_graph_phase_ms(hm_graph* g, const char* phase, double* ms) {
  hm::clear_error();
  if (!g || !phase || !ms) return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_phase_ms: null argument");
  if (std::strcmp(phase, "project") == 0) *ms = g->ms_project;
  else if (std::strcmp(phase, "triangles") == 0) *ms = g->ms_triangles;
  else if (std::strcmp(phase, "store") == 0) *ms = g->ms_upd_store;
  else return hm::fail(HM_ERR_BAD_INPUT, "hm_graph_phase_ms: unknown phase");
  return HM_OK;
}

}  // extern "C"
