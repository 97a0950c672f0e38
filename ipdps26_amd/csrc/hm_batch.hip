// Many-graph static count: one launch sequence counts a whole batch of device-resident graphs.
//
// The reference counts one graph per call on the default stream (dynamicHyperGraph/src/
// HMotifCount.cu:97-125); a count of a small graph there and in the one-graph path here is a
// dozen short launches and two host waits. A batch instead treats its graphs as one disjoint
// union: every hyperedge of every graph gets a global position p = base[g] + local id, and
//   k_batch_proj_count  one workgroup per position: projected degree and OpenAll into the
//                       graph's own accumulator (the body of the one-graph kernel, run on the
//                       graph's own CSR);
//   degree order        one radix sort of (graph << 32 | degree) keys: graph g's hyperedges land
//                       at positions base[g] .. base[g] + E_g - 1 in ascending degree, so a local
//                       rank is the sorted position minus base[g];
//   k_batch_proj_fill   upper rows of every graph into one row array (offsets by one scan);
//   k_batch_work        work estimate per centre; one descending sort orders the centres of all
//                       graphs together (largest first), so a long centre of one graph overlaps
//                       the short ones of the others;
//   k_batch_tri         the triangle body per centre, as one workgroup per centre or as a
//                       persistent grid that takes centres from a device work queue (an atomic
//                       counter), with waves per centre and LDS histogram copies as template
//                       parameters.
// The triangle form is picked per batch shape by an autotuner: the first batch of a shape runs
// every form, checks that they agree bin for bin, and keeps the fastest.
// Graphs whose mask table does not fit one LDS table are counted by the one-graph (tiled) path.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "hm_device.hpp"
#include "hm_graph.hpp"
#include "hm_internal.hpp"
#include "hm_static_dense.hpp"

namespace hm {

// One graph of a batch as the kernels see it.
struct BatchView {
  const uint32_t* eoff;
  const int32_t* enodes;
  const uint8_t* esize;
  const uint32_t* voff;
  const int32_t* vedges;
  int32_t E;
  int32_t base;  // global position of the graph's hyperedge 0
};

// gid[p] = graph of global position p (ebase ascending, every graph non-empty).
__global__ void k_batch_gid(int32_t n, const int32_t* __restrict__ ebase, int32_t Et,
                            int32_t* __restrict__ gid) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Et) return;
  int32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (ebase[mid] <= p) lo = mid;
    else hi = mid;
  }
  gid[p] = lo;
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_batch_proj_count(const BatchView* __restrict__ views,
                                                             const int32_t* __restrict__ gid,
                                                             uint32_t* __restrict__ deg,
                                                             unsigned long long* __restrict__ acc) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t p = blockIdx.x;
  const int32_t g = gid[p];
  const BatchView v = views[g];
  proj_count_item<MT>(reinterpret_cast<MT*>(smem), v.E, p - v.base, v.eoff, v.enodes, v.esize, v.voff,
                      v.vedges, deg + p, nullptr, acc + (size_t)g * kAccCount);
}

__global__ void k_batch_keys(int32_t Et, const int32_t* __restrict__ gid, const uint32_t* __restrict__ deg,
                             unsigned long long* __restrict__ keys, int32_t* __restrict__ ids) {
  const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Et) return;
  keys[p] = ((unsigned long long)gid[p] << 32) | deg[p];
  ids[p] = p;
}

// Sorted position q holds global position sperm[q] of graph gid[q]: local rank q - base.
__global__ void k_batch_ranks(int32_t Et, const BatchView* __restrict__ views,
                              const int32_t* __restrict__ gid,
                              const unsigned long long* __restrict__ skeys,
                              const int32_t* __restrict__ sperm, int32_t* __restrict__ rank,
                              int32_t* __restrict__ perm, uint8_t* __restrict__ szr,
                              unsigned long long* __restrict__ degr) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q == Et) degr[Et] = 0;
  if (q >= Et) return;
  const int32_t p = sperm[q];
  const int32_t b = views[gid[q]].base;
  rank[p] = q - b;
  perm[q] = p - b;
  szr[q] = views[gid[q]].esize[p - b];
  degr[q] = skeys[q] & 0xffffffffull;
}

template <typename MT>
__global__ __launch_bounds__(kBlock) void k_batch_proj_fill(
    const BatchView* __restrict__ views, const int32_t* __restrict__ gid,
    const int32_t* __restrict__ rank, const unsigned long long* __restrict__ roff,
    uint32_t* __restrict__ rlen, uint32_t* __restrict__ rows) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int32_t p = blockIdx.x;
  const BatchView v = views[gid[p]];
  proj_fill_item<MT>(reinterpret_cast<MT*>(smem), v.E, p - v.base, v.eoff, v.enodes, v.voff, v.vedges,
                     rank + v.base, roff + v.base, rlen + v.base, rows);
}

// Work estimate of the centre at sorted position q: its upper-row length plus the upper-row
// lengths of its upper neighbours.
__global__ __launch_bounds__(kBlock) void k_batch_work(int32_t Et, const BatchView* __restrict__ views,
                                                       const int32_t* __restrict__ gid,
                                                       const unsigned long long* __restrict__ roff,
                                                       const uint32_t* __restrict__ rlen,
                                                       const uint32_t* __restrict__ rows,
                                                       unsigned long long* __restrict__ work) {
  const int32_t q = blockIdx.x * kWavesPerBlock + wave_id();
  if (q >= Et) return;
  const int32_t b = views[gid[q]].base;
  const unsigned long long s0 = roff[q];
  const uint32_t n = rlen[q];
  unsigned long long s = 0;
  for (uint32_t j = lane_id(); j < n; j += kWave) s += rlen[b + (int32_t)(rows[s0 + j] >> kCntBits)];
  s = wave_sum(s);
  if (lane_id() == 0) work[q] = s + n;
}

// Triangles of the centres order[0 .. n_items). queue == nullptr: workgroup i takes centre
// order[i] (grid = n_items). Otherwise a persistent grid: each workgroup takes the next index
// from the counter *queue (zero at launch) until the indices run out, so every workgroup exits
// once n_items indices have been handed out.
template <typename MT, int NW, int NC>
__global__ __launch_bounds__(NW * kWave) void k_batch_tri(
    int32_t n_items, const int32_t* __restrict__ order, const BatchView* __restrict__ views,
    const int32_t* __restrict__ gid, const int32_t* __restrict__ rank,
    const int32_t* __restrict__ perm, const uint8_t* __restrict__ szr,
    const unsigned long long* __restrict__ roff, const uint32_t* __restrict__ rlen,
    const uint32_t* __restrict__ rows, unsigned long long* __restrict__ acc, int32_t* queue) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t s_item;
  MT* M = reinterpret_cast<MT*>(smem);
  for (int32_t it = (int32_t)blockIdx.x;;) {
    if (queue) {
      if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
      __syncthreads();
      it = s_item;
      __syncthreads();
    }
    if (it >= n_items) break;
    const int32_t q = order[it];
    const int32_t g = gid[q];
    const BatchView v = views[g];
    const int32_t b = v.base;
    tri_item<MT, NW, NC>(M, v.E, q - b, v.eoff, v.enodes, v.voff, v.vedges, rank + b, perm + b, szr + b,
                         roff + b, rlen + b, rows, acc + (size_t)g * kAccCount);
    if (!queue) break;
  }
}

// Work estimates saturate at 2^40 in the centre-order key (a centre's work is at most the
// square of its graph's hyperedge count, below 2^30 for one LDS table).
constexpr int kBatchWorkBits = 40;
constexpr unsigned long long kBatchWorkMax = (1ull << kBatchWorkBits) - 1;

__global__ void k_batch_zero(unsigned long long* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_batch_zero_i32(int32_t* p) { *p = 0; }

// caps[k] = first row slot of graph k (roff at its first sorted position), caps[n] = total.
__global__ void k_batch_cap_starts(int32_t n, const int32_t* __restrict__ ebase,
                                   const unsigned long long* __restrict__ roff,
                                   unsigned long long* __restrict__ caps) {
  const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= n) caps[k] = roff[ebase[k]];
}

// Sort key of the centre at sorted position q: its graph's chunk first (chunk 0 largest key),
// then its work estimate, so a descending sort runs chunk after chunk, most work first.
__global__ void k_batch_order_keys(int32_t Et, const int32_t* __restrict__ gid,
                                   const int32_t* __restrict__ chunk, int32_t n_chunks,
                                   const unsigned long long* __restrict__ work,
                                   unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Et) return;
  const unsigned long long w = work[q] < kBatchWorkMax ? work[q] : kBatchWorkMax;
  keys[q] = ((unsigned long long)(n_chunks - 1 - chunk[gid[q]]) << kBatchWorkBits) | w;
  vals[q] = q;
}

__global__ void k_batch_iota(int32_t* p, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// ---------------------------------------------------------------------------------------------
// Triangle forms and the autotuner
// ---------------------------------------------------------------------------------------------
struct TriForm {
  int waves, copies, persistent;
};
constexpr TriForm kTriForms[] = {{4, 8, 0}, {4, 16, 0}, {8, 8, 0}, {8, 16, 0},
                                 {4, 8, 1}, {4, 16, 1}, {8, 8, 1}, {8, 16, 1}};
constexpr int kNumTriForms = (int)(sizeof(kTriForms) / sizeof(kTriForms[0]));
// Centre orders: graph after graph (each graph's rows stay in the caches while its centres
// run), chunks of graphs holding up to 64 MB of rows, or all centres of the batch by work alone
// (the best balance, the largest working set).
constexpr int kNumOrders = 3;
// The autotuner searches only batches whose first triangle run takes at most this long (so a
// search costs at most ~11 such runs); longer batches keep graph order and the widest form,
// which measured within 1% of the best form on graphs of 1k-6k hyperedges.
constexpr float kTuneBudgetMs = 5.f;
constexpr unsigned long long kOrderChunkBytes[kNumOrders] = {0ull, 64ull << 20, ~0ull};
// LDS beside the dynamic table: triangle slot counters (243 slots x at most 16 copies x 4 B)
// plus the small tables.
constexpr size_t kBatchStaticLds = 20 * 1024;

}  // namespace hm

struct hm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hm::DevBuf views, ebase, gid, deg, keys, skeys, ids, sperm, rank, perm, szr, degr, roff, rlen, rows,
      work, order, ranks, acc, acc_tri, queue, caps, chunk, tmp;
  // pinned staging: descriptors and accumulators (stage), chunk ids of a centre order (h_chunk)
  hm::PinnedBuf stage, h_chunk;
  int forced_form = -1;  // -1: autotuned
  int forced_order = -1;
  // shape (mask width, log2 total hyperedges, log2 largest graph) -> fastest (order, form)
  std::map<std::tuple<int, int, int>, std::pair<int, int>> tuned;
  hm_batch_stats last{};
  int cu_count = 0;
};

namespace hm {
namespace {

int ceil_log2(int64_t x) {
  int k = 0;
  while (((int64_t)1 << k) < x) ++k;
  return k;
}

template <typename MT>
size_t table_bytes(int32_t E, bool with_sizes) {
  size_t v = (size_t)E * sizeof(MT) + (with_sizes ? (size_t)E : 0);
  return (v + 15) & ~(size_t)15;
}

// One triangle form: NW waves per centre, NC histogram copies per wave; persistent: a grid of
// (resident workgroups per CU) x CUs fed by the work counter, else one workgroup per centre.
template <typename MT, int NW, int NC>
hm_status launch_tri_form(hm_batch* b, bool persistent, int32_t Et, const BatchView* views, size_t lds,
                          unsigned long long* acc_tri) {
  const void* fn = (const void*)k_batch_tri<MT, NW, NC>;
  HM_TRY(allow_lds(fn, lds));
  int32_t* queue = nullptr;
  unsigned grid = (unsigned)Et;
  if (persistent) {
    int per_cu = 0;
    HM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NW * kWave, lds));
    const int64_t resident = (int64_t)std::max(per_cu, 1) * std::max(b->cu_count, 1);
    grid = (unsigned)std::min<int64_t>(Et, resident);
    HM_TRY(b->queue.ensure(sizeof(int32_t)));
    queue = b->queue.as<int32_t>();
    k_batch_zero_i32<<<1, 1, 0, b->stream>>>(queue);
  }
  k_batch_tri<MT, NW, NC><<<grid, NW * kWave, lds, b->stream>>>(
      Et, b->order.as<int32_t>(), views, b->gid.as<int32_t>(), b->rank.as<int32_t>(), b->perm.as<int32_t>(),
      b->szr.as<uint8_t>(), b->roff.as<unsigned long long>(), b->rlen.as<uint32_t>(), b->rows.as<uint32_t>(),
      acc_tri, queue);
  HM_HIP(hipGetLastError());
  return HM_OK;
}

template <typename MT>
hm_status launch_tri(hm_batch* b, int form, int32_t Et, const BatchView* views, size_t lds,
                     unsigned long long* acc_tri) {
  const TriForm f = kTriForms[form];
  const bool p = f.persistent != 0;
  if (f.waves == 4)
    return f.copies == 8 ? launch_tri_form<MT, 4, 8>(b, p, Et, views, lds, acc_tri)
                         : launch_tri_form<MT, 4, 16>(b, p, Et, views, lds, acc_tri);
  return f.copies == 8 ? launch_tri_form<MT, 8, 8>(b, p, Et, views, lds, acc_tri)
                       : launch_tri_form<MT, 8, 16>(b, p, Et, views, lds, acc_tri);
}

// Counts the graphs gs (all dense-eligible for MT) into out[k * 30 + bin].
template <typename MT>
hm_status count_batch(hm_batch* b, const std::vector<hm_graph*>& gs, std::vector<uint64_t*>& outs,
                      int* tuned_now) {
  const int32_t n = (int32_t)gs.size();
  hipStream_t st = b->stream;
  std::vector<int32_t> base((size_t)n + 1, 0);
  int32_t Emax = 0;
  for (int32_t k = 0; k < n; ++k) {
    base[(size_t)k + 1] = base[(size_t)k] + gs[(size_t)k]->csr.E;
    Emax = std::max(Emax, gs[(size_t)k]->csr.E);
  }
  const int32_t Et = base[(size_t)n];
  const size_t view_bytes = sizeof(BatchView) * (size_t)n;
  const size_t base_bytes = sizeof(int32_t) * ((size_t)n + 1);
  const size_t acc_bytes = sizeof(unsigned long long) * kAccCount * (size_t)n;
  HM_TRY(b->stage.ensure(std::max(view_bytes + base_bytes, std::max(2 * acc_bytes, 8 * ((size_t)n + 1)))));
  HM_TRY(b->h_chunk.ensure(sizeof(int32_t) * (size_t)n));
  BatchView* hv = b->stage.as<BatchView>();
  for (int32_t k = 0; k < n; ++k) {
    const DeviceCsr& c = gs[(size_t)k]->csr;
    hv[k] = BatchView{c.eoff.as<uint32_t>(), c.enodes.as<int32_t>(), c.esize.as<uint8_t>(),
                      c.voff.as<uint32_t>(), c.vedges.as<int32_t>(), c.E, base[(size_t)k]};
  }
  std::memcpy(b->stage.as<char>() + view_bytes, base.data(), base_bytes);

  HM_TRY(b->views.ensure(view_bytes));
  HM_TRY(b->ebase.ensure(base_bytes));
  HM_TRY(b->gid.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->deg.ensure(sizeof(uint32_t) * Et));
  HM_TRY(b->keys.ensure(sizeof(unsigned long long) * Et));
  HM_TRY(b->skeys.ensure(sizeof(unsigned long long) * Et));
  HM_TRY(b->ids.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->sperm.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->rank.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->perm.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->szr.ensure((size_t)Et));
  HM_TRY(b->degr.ensure(sizeof(unsigned long long) * ((size_t)Et + 1)));
  HM_TRY(b->roff.ensure(sizeof(unsigned long long) * ((size_t)Et + 1)));
  HM_TRY(b->rlen.ensure(sizeof(uint32_t) * Et));
  HM_TRY(b->work.ensure(sizeof(unsigned long long) * Et));
  HM_TRY(b->caps.ensure(sizeof(unsigned long long) * ((size_t)n + 1)));
  HM_TRY(b->chunk.ensure(sizeof(int32_t) * (size_t)n));
  HM_TRY(b->order.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->ranks.ensure(sizeof(int32_t) * Et));
  HM_TRY(b->acc.ensure(acc_bytes));
  HM_TRY(b->acc_tri.ensure(acc_bytes));

  const BatchView* views = b->views.as<BatchView>();
  const int32_t* gid = b->gid.as<int32_t>();
  unsigned long long* acc = b->acc.as<unsigned long long>();
  unsigned long long* acc_tri = b->acc_tri.as<unsigned long long>();
  const unsigned eb = (unsigned)((Et + kBlock - 1) / kBlock);
  const size_t lds_cnt = table_bytes<MT>(Emax, false);
  const size_t lds_tri = table_bytes<MT>(Emax, true);
  HM_TRY(allow_lds((const void*)k_batch_proj_count<MT>, lds_cnt));
  HM_TRY(allow_lds((const void*)k_batch_proj_fill<MT>, lds_cnt));

  HM_HIP(hipEventRecord(b->ev[0], st));
  HM_HIP(hipMemcpyAsync(b->views.p, b->stage.p, view_bytes, hipMemcpyHostToDevice, st));
  HM_HIP(hipMemcpyAsync(b->ebase.p, b->stage.as<char>() + view_bytes, base_bytes, hipMemcpyHostToDevice, st));
  k_batch_gid<<<eb, kBlock, 0, st>>>(n, b->ebase.as<int32_t>(), Et, b->gid.as<int32_t>());
  k_batch_zero<<<64, kBlock, 0, st>>>(acc, (int64_t)kAccCount * n);
  // 1. projected degrees + OpenAll, per graph
  k_batch_proj_count<MT><<<(unsigned)Et, kBlock, lds_cnt, st>>>(views, gid, b->deg.as<uint32_t>(), acc);
  HM_HIP(hipGetLastError());
  // 2. degree order within each graph: one sort of (graph << 32 | degree)
  k_batch_keys<<<eb, kBlock, 0, st>>>(Et, gid, b->deg.as<uint32_t>(), b->keys.as<unsigned long long>(),
                                      b->ids.as<int32_t>());
  const int key_bits = 32 + std::max(1, ceil_log2(n));
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, b->keys.as<unsigned long long>(),
                                            b->skeys.as<unsigned long long>(), b->ids.as<int32_t>(),
                                            b->sperm.as<int32_t>(), Et, 0, key_bits, st));
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, b->degr.as<unsigned long long>(),
                                          b->roff.as<unsigned long long>(), Et + 1, st));
  HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb3, b->keys.as<unsigned long long>(),
                                                      b->skeys.as<unsigned long long>(),
                                                      b->ranks.as<int32_t>(), b->order.as<int32_t>(),
                                                      Et, 0, 64, st));
  HM_TRY(b->tmp.ensure(std::max(tb1, std::max(tb2, tb3))));
  size_t tb = b->tmp.bytes;
  HM_HIP(hipcub::DeviceRadixSort::SortPairs(b->tmp.p, tb, b->keys.as<unsigned long long>(),
                                            b->skeys.as<unsigned long long>(), b->ids.as<int32_t>(),
                                            b->sperm.as<int32_t>(), Et, 0, key_bits, st));
  k_batch_ranks<<<(unsigned)((Et + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      Et, views, gid, b->skeys.as<unsigned long long>(), b->sperm.as<int32_t>(), b->rank.as<int32_t>(),
      b->perm.as<int32_t>(), b->szr.as<uint8_t>(), b->degr.as<unsigned long long>());
  // 3. row offsets (capacity d_a by sorted position)
  tb = b->tmp.bytes;
  HM_HIP(hipcub::DeviceScan::ExclusiveSum(b->tmp.p, tb, b->degr.as<unsigned long long>(),
                                          b->roff.as<unsigned long long>(), Et + 1, st));
  k_batch_cap_starts<<<(unsigned)((n + 1 + kBlock - 1) / kBlock), kBlock, 0, st>>>(
      n, b->ebase.as<int32_t>(), b->roff.as<unsigned long long>(), b->caps.as<unsigned long long>());
  HM_HIP(hipMemcpyAsync(b->stage.p, b->caps.p, sizeof(unsigned long long) * ((size_t)n + 1),
                        hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  const std::vector<unsigned long long> caps(b->stage.as<unsigned long long>(),
                                             b->stage.as<unsigned long long>() + n + 1);
  const unsigned long long total_cap = caps[(size_t)n];
  HM_TRY(b->rows.ensure(sizeof(uint32_t) * (size_t)std::max<unsigned long long>(total_cap, 1)));
  // 4. upper rows
  k_batch_proj_fill<MT><<<(unsigned)Et, kBlock, lds_cnt, st>>>(
      views, gid, b->rank.as<int32_t>(), b->roff.as<unsigned long long>(), b->rlen.as<uint32_t>(),
      b->rows.as<uint32_t>());
  HM_HIP(hipGetLastError());
  // 5. work estimate per centre
  k_batch_work<<<(unsigned)((Et + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, st>>>(
      Et, views, gid, b->roff.as<unsigned long long>(), b->rlen.as<uint32_t>(), b->rows.as<uint32_t>(),
      b->work.as<unsigned long long>());
  HM_HIP(hipEventRecord(b->ev[1], st));

  // centre order `oi`: chunk ids per graph (rows bytes up to the chunk budget), then one sort
  auto make_order = [&](int oi) -> hm_status {
    int32_t* hc = b->h_chunk.as<int32_t>();
    int32_t c = 0;
    unsigned long long run = 0;
    for (int32_t k = 0; k < n; ++k) {
      const unsigned long long bytes = 4 * (caps[(size_t)k + 1] - caps[(size_t)k]);
      if (k > 0 && (oi == 0 || run + bytes > kOrderChunkBytes[oi])) {
        ++c;
        run = 0;
      }
      run += bytes;
      hc[k] = c;
    }
    const int32_t n_chunks = c + 1;
    HM_HIP(hipMemcpyAsync(b->chunk.p, hc, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st));
    k_batch_order_keys<<<eb, kBlock, 0, st>>>(Et, gid, b->chunk.as<int32_t>(), n_chunks,
                                              b->work.as<unsigned long long>(),
                                              b->keys.as<unsigned long long>(), b->ranks.as<int32_t>());
    size_t tbs = b->tmp.bytes;
    HM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(
        b->tmp.p, tbs, b->keys.as<unsigned long long>(), b->skeys.as<unsigned long long>(),
        b->ranks.as<int32_t>(), b->order.as<int32_t>(), Et, 0,
        std::min(64, kBatchWorkBits + std::max(1, ceil_log2(n_chunks))), st));
    return HM_OK;
  };

  // 6. triangles with the forced configuration, the one tuned for this shape, or tune now
  unsigned long long* h_tri = b->stage.as<unsigned long long>() + kAccCount * (size_t)n;
  const int64_t acc_words = (int64_t)kAccCount * n;
  auto run_tri = [&](int form, float* ms) -> hm_status {
    k_batch_zero<<<64, kBlock, 0, st>>>(acc_tri, acc_words);
    HM_HIP(hipEventRecord(b->ev[2], st));
    HM_TRY(launch_tri<MT>(b, form, Et, views, lds_tri, acc_tri));
    HM_HIP(hipEventRecord(b->ev[3], st));
    HM_HIP(hipMemcpyAsync(h_tri, acc_tri, acc_bytes, hipMemcpyDeviceToHost, st));
    HM_HIP(hipStreamSynchronize(st));
    HM_HIP(hipEventElapsedTime(ms, b->ev[2], b->ev[3]));
    return HM_OK;
  };
  const auto key = std::make_tuple((int)sizeof(MT), ceil_log2(Et), ceil_log2(Emax));
  int oi = b->forced_order, form = b->forced_form;
  if (form < 0) {
    auto it = b->tuned.find(key);
    if (it != b->tuned.end()) std::tie(oi, form) = it->second;
  }
  float ms = 0.f;
  if (form >= 0) {
    HM_TRY(make_order(oi));
    HM_TRY(run_tri(form, &ms));
  } else {
    // coordinate search: the orders with the widest form, then every form with the best order;
    // every run must give the same bins
    constexpr int kProbeForm = 3;  // 8 waves, 16 copies, one workgroup per centre
    std::vector<unsigned long long> first;
    auto check = [&](int tag) -> hm_status {
      if (first.empty()) {
        first.assign(h_tri, h_tri + acc_words);
      } else if (std::memcmp(first.data(), h_tri, acc_bytes) != 0) {
        set_error("hm_batch: autotuner run %d disagrees with the first run", tag);
        return HM_ERR_DEVICE;
      }
      return HM_OK;
    };
    float best = 0.f;
    for (int o = 0; o < kNumOrders; ++o) {
      float t = 0.f;
      HM_TRY(make_order(o));
      HM_TRY(run_tri(kProbeForm, &t));
      HM_TRY(check(o));
      if (o == 0 || t < best) {
        best = t;
        oi = o;
      }
      if (o == 0 && t > kTuneBudgetMs) break;  // a long batch: keep graph order and the widest form
    }
    form = kProbeForm;
    if (oi != kNumOrders - 1) HM_TRY(make_order(oi));  // the last order tried is in place already
    for (int f = 0; f < kNumTriForms && best <= kTuneBudgetMs; ++f) {
      if (f == kProbeForm) continue;
      float t = 0.f;
      HM_TRY(run_tri(f, &t));
      HM_TRY(check(kNumOrders + f));
      if (t < best) {
        best = t;
        form = f;
      }
    }
    b->tuned[key] = {oi, form};
    ms = best;
    *tuned_now = 1;
  }
  b->last.ms_triangles += ms;
  unsigned long long* h_acc = b->stage.as<unsigned long long>();
  HM_HIP(hipMemcpyAsync(h_acc, acc, acc_bytes, hipMemcpyDeviceToHost, st));
  HM_HIP(hipStreamSynchronize(st));
  float ms_pre = 0.f;
  HM_HIP(hipEventElapsedTime(&ms_pre, b->ev[0], b->ev[1]));
  b->last.ms_prepare += ms_pre;
  for (int32_t k = 0; k < n; ++k)
    for (int i = 0; i < HM_STATIC_BINS; ++i)
      outs[(size_t)k][i] = h_acc[(size_t)k * kAccCount + kAccHist + i] +
                           h_tri[(size_t)k * kAccCount + kAccHist + i];
  b->last.waves = kTriForms[form].waves;
  b->last.copies = kTriForms[form].copies;
  b->last.persistent = kTriForms[form].persistent;
  b->last.order = oi;
  b->last.hyperedges += Et;
  return HM_OK;
}

}  // namespace
}  // namespace hm

extern "C" {

hm_status hm_batch_create(hm_batch** out) {
  hm::clear_error();
  if (!out) return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_create: null out");
  *out = nullptr;
  hm_batch* b = new (std::nothrow) hm_batch();
  if (!b) return hm::fail(HM_ERR_OOM, "hm_batch_create: host allocation failed");
  b->device = hm::current_device();
  auto cleanup = [b](hm_status st) {
    for (hipEvent_t& e : b->ev)
      if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return st;
  };
  if (hipSetDevice(b->device) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hm_batch_create: no device"));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(hm::fail(HM_ERR_DEVICE, "hm_batch_create: hipStreamCreate failed"));
  for (hipEvent_t& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(hm::fail(HM_ERR_DEVICE, "hm_batch_create: hipEventCreate failed"));
  if (hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess)
    b->cu_count = 1;
  *out = b;
  return HM_OK;
}

hm_status hm_batch_destroy(hm_batch* b) {
  if (!b) return HM_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  for (hm::DevBuf* d : {&b->views, &b->ebase, &b->gid, &b->deg, &b->keys, &b->skeys, &b->ids, &b->sperm,
                        &b->rank, &b->perm, &b->szr, &b->degr, &b->roff, &b->rlen, &b->rows, &b->work,
                        &b->order, &b->ranks, &b->acc, &b->acc_tri, &b->queue, &b->caps, &b->chunk,
                        &b->tmp})
    d->release();
  b->stage.release();
  b->h_chunk.release();
  for (hipEvent_t& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return HM_OK;
}

hm_status hm_batch_set_form(hm_batch* b, int32_t waves, int32_t copies, int32_t persistent,
                            int32_t order) {
  hm::clear_error();
  if (!b) return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_set_form: null batch");
  if (waves == 0) {
    b->forced_form = b->forced_order = -1;
    return HM_OK;
  }
  if (order < 0 || order >= hm::kNumOrders)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_set_form: order must be 0, 1 or 2");
  for (int f = 0; f < hm::kNumTriForms; ++f)
    if (hm::kTriForms[f].waves == waves && hm::kTriForms[f].copies == copies &&
        hm::kTriForms[f].persistent == (persistent ? 1 : 0)) {
      b->forced_form = f;
      b->forced_order = order;
      return HM_OK;
    }
  return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_set_form: waves must be 4 or 8 and copies 8 or 16");
}

hm_status hm_batch_count_static(hm_batch* b, hm_graph* const* graphs, int32_t n, uint64_t* out) {
  hm::clear_error();
  if (!b || n < 0 || (n > 0 && (!graphs || !out)))
    return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_count_static: bad argument");
  for (int32_t k = 0; k < n; ++k) {
    if (!graphs[k]) return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_count_static: null graph");
    if (graphs[k]->device != b->device)
      return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_count_static: graph on another device than the batch");
    if (graphs[k]->csr.E >= 3 && graphs[k]->csr.max_size > HM_MAX_EDGE_SIZE)
      return hm::fail(HM_ERR_UNSUPPORTED, "hm_batch_count_static: hyperedge larger than 64 nodes");
    if ((int64_t)graphs[k]->csr.E >= hm::kMaxIds)
      return hm::fail(HM_ERR_UNSUPPORTED, "hm_batch_count_static: more than 2^22 - 1 hyperedges");
  }
  HM_HIP(hipSetDevice(b->device));
  b->last = hm_batch_stats{};
  int lim = 0;
  if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) != hipSuccess) lim = 65536;
  // batches by mask width, each cut so that its global positions stay below 2^30
  constexpr int64_t kMaxBatchEdges = (int64_t)1 << 30;
  std::vector<hm_graph*> part[2];
  std::vector<uint64_t*> outs[2];
  std::vector<int32_t> alone;
  int tuned = 0;
  auto flush = [&](int w) -> hm_status {
    if (part[w].empty()) return HM_OK;
    hm_status st = w ? hm::count_batch<hm::mask64_t>(b, part[w], outs[w], &tuned)
                     : hm::count_batch<uint32_t>(b, part[w], outs[w], &tuned);
    b->last.graphs_batched += (int32_t)part[w].size();
    part[w].clear();
    outs[w].clear();
    return st;
  };
  int64_t edges[2] = {0, 0};
  for (int32_t k = 0; k < n; ++k) {
    hm_graph* g = graphs[k];
    uint64_t* o = out + (size_t)k * HM_STATIC_BINS;
    const int32_t E = g->csr.E;
    if (E < 3) {
      std::memset(o, 0, sizeof(uint64_t) * HM_STATIC_BINS);
      continue;
    }
    const int w = g->csr.max_size > 32 ? 1 : 0;
    const size_t need = (w ? hm::table_bytes<hm::mask64_t>(E, true) : hm::table_bytes<uint32_t>(E, true)) +
                        hm::kBatchStaticLds;
    if (need > (size_t)lim) {
      alone.push_back(k);
      continue;
    }
    if (edges[w] + E > kMaxBatchEdges) {
      HM_TRY(flush(w));
      edges[w] = 0;
    }
    part[w].push_back(g);
    outs[w].push_back(o);
    edges[w] += E;
  }
  HM_TRY(flush(0));
  HM_TRY(flush(1));
  for (int32_t k : alone) {  // larger than one LDS table: the one-graph (tiled) count
    HM_TRY(hm_count_static(graphs[k], out + (size_t)k * HM_STATIC_BINS));
    ++b->last.graphs_alone;
  }
  b->last.tuned = tuned;
  return HM_OK;
}

hm_status hm_batch_get_stats(hm_batch* b, hm_batch_stats* out) {
  hm::clear_error();
  if (!b || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_batch_get_stats: null argument");
  *out = b->last;
  return HM_OK;
}

}  // extern "C"
