// Three-snapshot "strict-increasing" count: the semantics of the reference's temporal GPU path
// computeTemporalMotifCountsStrictInc (temporalDynamicHMotif/include/temporal_count.hpp:8-9,
// temporalDynamicHMotif/src/temporal_count.cu:91-217, temporalStrictIncAdjKernel).
//
// Three snapshots (older O, middle M, newest N) of a hypergraph, each a list of rows indexed by
// hyperedge id. A triple is (i, j, k) with i < j < k, row i of O, row j of M and row k of N
// pairwise overlapping; it is classified into the 30 static bins from the three rows (sizes,
// pairwise overlaps, common part). This is not THyMe's temporal motif semantics — it is what the
// reference's temporal-window driver computes — and it is kept for callers of that interface.
// Difference: the reference keeps at most 1,024 candidates per side of j and silently drops the
// rest (temporal_count.cu:79-89); here every candidate counts.
//
// Kernels:
//   k_xdeg / k_xfill  cross rows X(i) of O against N: (k << 7 | C_ik) for k > i + 1 with
//                     O[i] ∩ N[k] ≠ ∅, ascending (a dense LDS table over N's ids per i)
//   k_si              one workgroup per middle row j: tables MO[i] = j's nodes in O[i] and
//                     MN[k] = j's nodes in N[k]; for every i < j with MO[i] ≠ 0, the part of X(i)
//                     after j: MN[k] ≠ 0 closes the triangle, C_ij = popc(MO[i]),
//                     C_jk = popc(MN[k]), g = popc(MO[i] & MN[k]).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_tables.hpp"

namespace hm {

__constant__ int8_t c_si_bin[128] = HM_STATIC_BIN_TABLE;

namespace {

// One snapshot on the device: H2V and V2H over a vertex range shared by the three snapshots.
struct Layer {
  int32_t E = 0;
  int64_t I = 0;
  DevBuf eoff, enodes, esize, voff, vedges;
  ~Layer() {
    for (DevBuf* b : {&eoff, &enodes, &esize, &voff, &vedges}) b->release();
  }
};

hm_status upload_layer(Layer& L, const std::vector<int64_t>& off, const std::vector<int32_t>& nodes,
                       int32_t E, int32_t V, hipStream_t st) {
  L.E = E;
  L.I = (int64_t)nodes.size();
  std::vector<uint32_t> eoff((size_t)E + 1), voff((size_t)V + 1, 0);
  std::vector<uint8_t> esz((size_t)std::max(E, 1), 0);
  for (int32_t e = 0; e <= E; ++e) eoff[(size_t)e] = (uint32_t)off[(size_t)e];
  for (int32_t e = 0; e < E; ++e) esz[(size_t)e] = (uint8_t)(off[(size_t)e + 1] - off[(size_t)e]);
  for (int32_t v : nodes) voff[(size_t)v + 1]++;
  for (int32_t v = 0; v < V; ++v) voff[(size_t)v + 1] += voff[(size_t)v];
  std::vector<int32_t> vedges((size_t)std::max<int64_t>(L.I, 1));
  std::vector<uint32_t> fill(voff.begin(), voff.end() - 1);
  for (int32_t e = 0; e < E; ++e)
    for (int64_t q = off[(size_t)e]; q < off[(size_t)e + 1]; ++q)
      vedges[fill[(size_t)nodes[(size_t)q]]++] = e;
  HM_TRY(L.eoff.ensure(sizeof(uint32_t) * ((size_t)E + 1)));
  HM_TRY(L.enodes.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(L.I, 1)));
  HM_TRY(L.esize.ensure((size_t)std::max(E, 1)));
  HM_TRY(L.voff.ensure(sizeof(uint32_t) * ((size_t)V + 1)));
  HM_TRY(L.vedges.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(L.I, 1)));
  HM_HIP(hipMemcpyAsync(L.eoff.p, eoff.data(), sizeof(uint32_t) * eoff.size(), hipMemcpyHostToDevice, st));
  if (L.I > 0) {
    HM_HIP(hipMemcpyAsync(L.enodes.p, nodes.data(), sizeof(int32_t) * (size_t)L.I, hipMemcpyHostToDevice, st));
    HM_HIP(hipMemcpyAsync(L.vedges.p, vedges.data(), sizeof(int32_t) * (size_t)L.I, hipMemcpyHostToDevice, st));
  }
  HM_HIP(hipMemcpyAsync(L.esize.p, esz.data(), esz.size(), hipMemcpyHostToDevice, st));
  HM_HIP(hipMemcpyAsync(L.voff.p, voff.data(), sizeof(uint32_t) * voff.size(), hipMemcpyHostToDevice, st));
  HM_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  return HM_OK;
}

}  // namespace

// Number of k > i + 1 whose newest row overlaps older row i.
template <typename MT>
__global__ __launch_bounds__(kBlock) void k_xdeg(int32_t EN, const uint32_t* __restrict__ eoffO,
                                                 const int32_t* __restrict__ enodesO,
                                                 const uint32_t* __restrict__ voffN,
                                                 const int32_t* __restrict__ vedgesN,
                                                 unsigned long long* __restrict__ deg) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t red[kWavesPerBlock];
  const int32_t i = blockIdx.x;
  build_masks<MT>(M, EN, i, eoffO, enodesO, voffN, vedgesN, nullptr);
  uint32_t d = 0;
  for (int32_t k = i + 2 + (int32_t)threadIdx.x; k < EN; k += blockDim.x) d += M[k] != (MT)0;
  d = wave_sum(d);
  if (lane_id() == 0) red[wave_id()] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) t += red[q];
    deg[i] = t;
  }
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_xfill(int32_t EN, const uint32_t* __restrict__ eoffO,
                                                  const int32_t* __restrict__ enodesO,
                                                  const uint32_t* __restrict__ voffN,
                                                  const int32_t* __restrict__ vedgesN,
                                                  const unsigned long long* __restrict__ xoff,
                                                  uint32_t* __restrict__ xrows) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* M = reinterpret_cast<MT*>(smem);
  __shared__ uint32_t wcnt[kWavesPerBlock];
  const int32_t i = blockIdx.x;
  build_masks<MT>(M, EN, i, eoffO, enodesO, voffN, vedgesN, nullptr);
  const int lane = lane_id(), w = wave_id();
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t* out = xrows + xoff[i];
  uint32_t running = 0;
  for (int32_t k0 = i + 2; k0 < EN; k0 += blockDim.x) {
    const int32_t k = k0 + (int32_t)threadIdx.x;
    const MT m = k < EN ? M[k] : (MT)0;
    const unsigned long long bal = __ballot(m != (MT)0);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < kWavesPerBlock; ++q) {
      woff += q < w ? wcnt[q] : 0;
      tot += wcnt[q];
    }
    if (m) out[running + woff + (uint32_t)__popcll(bal & below)] = ((uint32_t)k << kCntBits) | (uint32_t)popc(m);
    running += tot;
    __syncthreads();
  }
}

constexpr int kSiWaves = 8;
constexpr int kSiCopies = 8;

template <typename MT, int NW>
__global__ __launch_bounds__(NW * kWave) void k_si(
    int32_t EO, int32_t EN, const uint32_t* __restrict__ eoffM, const int32_t* __restrict__ enodesM,
    const uint8_t* __restrict__ esizeO, const uint8_t* __restrict__ esizeN,
    const uint32_t* __restrict__ voffO, const int32_t* __restrict__ vedgesO,
    const uint32_t* __restrict__ voffN, const int32_t* __restrict__ vedgesN,
    const unsigned long long* __restrict__ xoff, const unsigned long long* __restrict__ xdeg,
    const uint32_t* __restrict__ xrows, unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  MT* MO = reinterpret_cast<MT*>(smem);
  MT* MN = MO + (((size_t)EO + 3) & ~(size_t)3);
  __shared__ int32_t h[NW][32][kSiCopies];
  __shared__ int8_t tbl[128];
  const int32_t j = blockIdx.x;
  const int sj = (int)(eoffM[j + 1] - eoffM[j]);
  if (sj == 0 || j == 0) return;  // no older row precedes row 0
  for (int k = threadIdx.x; k < NW * 32 * kSiCopies; k += blockDim.x) (&h[0][0][0])[k] = 0;
  if (threadIdx.x < 128) tbl[threadIdx.x] = c_si_bin[threadIdx.x];
  build_masks<MT>(MO, EO, j, eoffM, enodesM, voffO, vedgesO, nullptr);  // every O id may appear
  build_masks<MT>(MN, EN, j, eoffM, enodesM, voffN, vedgesN, nullptr);
  const int lane = lane_id(), w = wave_id();
  int32_t* hc = &h[w][0][lane & (kSiCopies - 1)];
  const int32_t iend = min(EO, j);
  for (int32_t i0 = w * kWave; i0 < iend; i0 += NW * kWave) {
    const int32_t i = i0 + lane;
    const MT mi = i < iend ? MO[i] : (MT)0;
    unsigned long long cand = __ballot(mi != (MT)0);
    while (cand) {
      const int leader = __ffsll((long long)cand) - 1;
      cand &= cand - 1;
      const int32_t ci = __shfl(i, leader);
      const MT mo = __shfl(mi, leader);
      const int si = esizeO[ci];
      const int cij = popc(mo);
      const uint32_t* row = xrows + xoff[ci];
      const uint32_t n = (uint32_t)xdeg[ci];
      // first entry with k > j (rows ascending)
      uint32_t lo = 0, hi = n;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((int32_t)(row[mid] >> kCntBits) <= j) lo = mid + 1;
        else hi = mid;
      }
      for (uint32_t q = lo + lane; q < n; q += kWave) {
        const uint32_t e = row[q];
        const int32_t k = (int32_t)(e >> kCntBits);
        const MT mn = MN[k];
        if (mn) {
          const int cik = (int)(e & kCntMask);
          atomicAdd(&hc[tbl[venn_id(si, sj, (int)esizeN[k], cij, popc(mn), cik, popc((MT)(mo & mn)))] * kSiCopies], 1);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < kStaticBins) {
    long long s = 0;
    for (int q = 0; q < NW; ++q)
      for (int c = 0; c < kSiCopies; ++c) s += h[q][threadIdx.x][c];
    if (s) atomicAdd(&acc[threadIdx.x], (unsigned long long)s);
  }
}

__global__ void k_si_zero(unsigned long long* p, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

template <typename MT>
static hm_status strict_inc_impl(Layer& O, Layer& Mi, Layer& N, hipStream_t st, int lim,
                                 unsigned long long* h_acc) {
  const int32_t EO = O.E, EM = Mi.E, EN = N.E;
  const size_t lds_x = (((size_t)EN * sizeof(MT)) + 15) & ~(size_t)15;
  const size_t lds_si = ((((size_t)EO + 3) & ~(size_t)3) * sizeof(MT) + (size_t)EN * sizeof(MT) + 15) & ~(size_t)15;
  const size_t static_si = sizeof(int32_t) * kSiWaves * 32 * kSiCopies + 1024;
  if (lds_x > (size_t)lim || lds_si + static_si > (size_t)lim) {
    set_error("strict-inc count: %d older + %d newest rows exceed the LDS tables", EO, EN);
    return HM_ERR_UNSUPPORTED;
  }
  HM_TRY(allow_lds((const void*)k_xdeg<MT>, lds_x));
  HM_TRY(allow_lds((const void*)k_xfill<MT>, lds_x));
  HM_TRY(allow_lds((const void*)k_si<MT, kSiWaves>, lds_si));
  DevBuf xdeg, xoff, xrows, tmp, acc;
  struct Release {
    DevBuf* b[5];
    ~Release() { for (DevBuf* x : b) x->release(); }
  } rel{{&xdeg, &xoff, &xrows, &tmp, &acc}};
  HM_TRY(xdeg.ensure(sizeof(unsigned long long) * ((size_t)EO + 1)));
  HM_TRY(xoff.ensure(sizeof(unsigned long long) * ((size_t)EO + 1)));
  HM_TRY(acc.ensure(sizeof(unsigned long long) * 32));
  k_si_zero<<<1, 64, 0, st>>>(acc.as<unsigned long long>(), 32);
  k_si_zero<<<1, 64, 0, st>>>(xdeg.as<unsigned long long>() + EO, 1);
  if (EO > 0)
    k_xdeg<MT><<<EO, kBlock, lds_x, st>>>(EN, O.eoff.as<uint32_t>(), O.enodes.as<int32_t>(),
                                          N.voff.as<uint32_t>(), N.vedges.as<int32_t>(),
                                          xdeg.as<unsigned long long>());
  HM_HIP(hipGetLastError());
  size_t tb = 0;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, xdeg.as<unsigned long long>(),
                                          xoff.as<unsigned long long>(), EO + 1, st));
  HM_TRY(tmp.ensure(tb));
  tb = tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, xdeg.as<unsigned long long>(),
                                          xoff.as<unsigned long long>(), EO + 1, st));
  unsigned long long total = 0;
  HM_HIP(hipMemcpyAsync(&total, xoff.as<unsigned long long>() + EO, 8, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  HM_TRY(xrows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total, 1)));
  if (EO > 0)
    k_xfill<MT><<<EO, kBlock, lds_x, st>>>(EN, O.eoff.as<uint32_t>(), O.enodes.as<int32_t>(),
                                           N.voff.as<uint32_t>(), N.vedges.as<int32_t>(),
                                           xoff.as<unsigned long long>(), xrows.as<uint32_t>());
  HM_HIP(hipGetLastError());
  if (EM > 0)
    k_si<MT, kSiWaves><<<EM, kSiWaves * kWave, lds_si, st>>>(
        EO, EN, Mi.eoff.as<uint32_t>(), Mi.enodes.as<int32_t>(), O.esize.as<uint8_t>(),
        N.esize.as<uint8_t>(), O.voff.as<uint32_t>(), O.vedges.as<int32_t>(), N.voff.as<uint32_t>(),
        N.vedges.as<int32_t>(), xoff.as<unsigned long long>(), xdeg.as<unsigned long long>(),
        xrows.as<uint32_t>(), acc.as<unsigned long long>());
  HM_HIP(hipGetLastError());
  HM_HIP(hipMemcpyAsync(h_acc, acc.p, sizeof(unsigned long long) * 32, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  return HM_OK;
}

}  // namespace hm

extern "C" hm_status hm_count_strict_inc(const int64_t* const offsets[3], const int32_t* const nodes[3],
                                         const int32_t num_rows[3], uint64_t out[HM_STATIC_BINS]) {
  using namespace hm;
  clear_error();
  if (!offsets || !nodes || !num_rows || !out) return fail(HM_ERR_BAD_INPUT, "hm_count_strict_inc: null argument");
  std::vector<int64_t> off[3];
  std::vector<int32_t> nod[3];
  int32_t vmax = -1, maxsz = 0;
  for (int l = 0; l < 3; ++l) {
    if (num_rows[l] < 0) return fail(HM_ERR_BAD_INPUT, "hm_count_strict_inc: negative row count");
    if ((int64_t)num_rows[l] >= kMaxPackedId) return fail(HM_ERR_UNSUPPORTED, "hm_count_strict_inc: more than 2^25 rows");
    int32_t vm = -1;
    HM_TRY(prepare_host_rows(offsets[l], nodes[l], num_rows[l], &off[l], &nod[l], &vm));
    if ((int64_t)nod[l].size() >= ((int64_t)1 << 32)) return fail(HM_ERR_OVER_BUDGET, "hm_count_strict_inc: more than 2^32 incidences");
    vmax = std::max(vmax, vm);
    for (int32_t e = 0; e < num_rows[l]; ++e)
      maxsz = std::max<int32_t>(maxsz, (int32_t)(off[l][(size_t)e + 1] - off[l][(size_t)e]));
  }
  if (maxsz > HM_MAX_EDGE_SIZE) return fail(HM_ERR_UNSUPPORTED, "hm_count_strict_inc: row larger than 64 nodes");
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = 0;
  const int32_t V = vmax + 1;
  const int dev = current_device();
  HM_HIP(hipSetDevice(dev));
  int lim = 65536;
  (void)hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  hipStream_t st = nullptr;
  HM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{st};
  Layer L[3];
  for (int l = 0; l < 3; ++l) HM_TRY(upload_layer(L[l], off[l], nod[l], num_rows[l], std::max(V, 1), st));
  unsigned long long acc[32] = {0};
  HM_TRY(maxsz > 32 ? strict_inc_impl<mask64_t>(L[0], L[1], L[2], st, lim, acc)
                    : strict_inc_impl<uint32_t>(L[0], L[1], L[2], st, lim, acc));
  for (int k = 0; k < HM_STATIC_BINS; ++k) out[k] = acc[k];
  return HM_OK;
}
