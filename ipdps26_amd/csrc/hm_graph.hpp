// Host-side definition of the hm_graph handle and device buffer helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "hm_internal.hpp"

namespace hm {

// A device allocation that grows on demand and is reused across calls.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  // Regrowth adds 50% headroom: graphs grow a little with every update batch, and hipFree
  // synchronises the whole device, which would serialise concurrent requests.
  hm_status ensure(size_t need) {
    if (need <= bytes && p) return HM_OK;
    if (alloc_forbidden()) {
      set_error("a buffer of %zu bytes is needed; this entry point allocates nothing and its "
                "scratch holds %zu", need, bytes);
      return HM_ERR_OVER_BUDGET;
    }
    AllocGuard guard;
    size_t want = need ? need : 16;
    if (p) {
      want = std::max(want, bytes + bytes / 2);
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? HM_ERR_OOM : HM_ERR_DEVICE;
    }
    bytes = want;
    return HM_OK;
  }
  void release() {
    if (p) {
      AllocGuard guard;
      (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
};

// Pinned host memory that grows on demand (staging for asynchronous copies).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  // The new block is allocated before the old one is freed, so a failed growth leaves the old
  // block (and its address) in place. Callers drain every copy from the old block first.
  hm_status ensure(size_t need) {
    if (need <= bytes && p) return HM_OK;
    if (alloc_forbidden()) {
      set_error("a pinned buffer of %zu bytes is needed; this entry point allocates nothing", need);
      return HM_ERR_OVER_BUDGET;
    }
    AllocGuard guard;
    const size_t want = std::max(need, (size_t)1 << 16);
    void* np = nullptr;
    if (hipHostMalloc(&np, want, hipHostMallocDefault) != hipSuccess || !np) {
      set_error("hipHostMalloc(%zu bytes) failed", want);
      return HM_ERR_OOM;
    }
    if (p) (void)hipHostFree(p);
    p = np;
    bytes = want;
    return HM_OK;
  }
  void release() {
    if (p) {
      AllocGuard guard;
      (void)hipHostFree(p);
    }
    p = nullptr;
    bytes = 0;
  }
};

// Flat, sorted, HBM-resident incidence arrays of one hypergraph.
//   H2V: eoff[E+1] (uint32), enodes[I] (int32, ascending per row), esize[E] (uint8, 0 = deleted)
//   V2H: voff[V+1] (uint32), vedges[I] (int32, ascending per row)
struct DeviceCsr {
  int32_t E = 0, V = 0;
  int64_t I = 0;
  int32_t max_size = 0;
  DevBuf eoff, enodes, esize, voff, vedges;
  void release() {
    eoff.release();
    enodes.release();
    esize.release();
    voff.release();
    vedges.release();
  }
};

// Scratch for one static count (reused across calls on the same handle).
struct CountScratch {
  DevBuf deg;      // uint32[E]   projected degree d_a (by id)
  DevBuf ids;      // int32[E]    iota, radix-sort values
  DevBuf perm;     // int32[E]    rank -> id (ascending degree, ties by id)
  DevBuf rank;     // int32[E]    id -> rank
  DevBuf degr;     // uint32[E+1] degree by rank (sorted keys), then scanned into roff
  DevBuf roff;     // uint64[E+1] row offsets by rank (capacity = d_a)
  DevBuf rlen;     // uint32[E]   upper-row length d+ by rank
  DevBuf rows;     // uint32[P]   packed upper rows (rank << 7 | overlap)
  DevBuf szr;      // uint8[E]    hyperedge size by rank
  DevBuf work;     // uint64[E]   triangle-kernel work estimate by rank
  DevBuf work2;    // uint64[E]   sorted work (discarded)
  DevBuf order;    // int32[E]    centre ranks, most work first
  DevBuf ranks;    // int32[E]    iota for the work sort
  DevBuf dlow;     // uint32[E]   neighbours with a smaller id, per a
  DevBuf acc;      // uint64[64]  histogram (30 bins) + statistics
  DevBuf acc2;     // uint64[64]  triangle-form autotuning runs
  DevBuf tmp;      // hipcub temporary storage
  // tiled path only
  DevBuf vkeys;    // uint64[I]   (node << 32 | rank) of every incidence, then sorted
  DevBuf vkeys2;   // uint64[I]
  DevBuf vrank;    // int32[I]    V2H relabelled to ranks, ascending per node
  DevBuf uoff;     // uint64[E+1] exclusive scan of the upper-row lengths (state offsets)
  DevBuf cur;      // uint32[U]   read position in b's row, per upper-row entry (a, b)
  DevBuf pm;       // mask[U]     a's nodes contained in b, per upper-row entry (a, b)
  void release() {
    for (DevBuf* b : {&deg, &ids, &perm, &rank, &degr, &roff, &rlen, &rows, &szr, &work, &work2,
                      &order, &ranks, &dlow, &acc, &acc2, &tmp, &vkeys, &vkeys2, &vrank, &uoff, &cur, &pm})
      b->release();
  }
};

// Scratch of the hub-aware static count (hm_hub.hip), reused across calls.
struct HubScratch {
  DevBuf hubidx;   // int8[V]     hub index of a node (-1 = light)
  DevBuf hmask;    // uint32[E]   hub pattern of a hyperedge
  DevBuf lsz;      // uint8[E]    light-part size
  DevBuf ckey;     // uint64[E]   class keys (hub pattern << 1 | [light part non-empty])
  DevBuf ckey2;    // uint64[E]   sorted class keys
  DevBuf ukey;     // uint64[K]   distinct class keys
  DevBuf ucnt;     // uint32[K]   hyperedges per class
  DevBuf nruns;    // uint64[4]   class count, light-pair count, pair-type count
  DevBuf hmr;      // uint32[E]   hub pattern by light rank
  DevBuf lszr;     // uint8[E]    light size by light rank
  DevBuf types;    // uint64[N]   pair-type key per (centre, neighbour key) entry
  DevBuf tcraw;    // uint64[N]   light pairs of the entry
  DevBuf types2;   // uint64[N]   sorted pair-type keys
  DevBuf tcsrt;    // uint64[N]   their counts
  DevBuf tkey;     // uint64[T]   distinct pair types
  DevBuf tcnt;     // uint64[T]   light pairs per type
  DevBuf rkey;     // uint32[E]   light-order sort keys
  DevBuf first;    // uint8[E]    hyperedges ranked first (an update's changed ids)
  DevBuf firstids; // int32[n]    their ids
  DevBuf accd;     // uint64[40]  (D)'s bins
  DevBuf accr;     // uint64[40]  (C) of one range of centres
  DevBuf big;      // int32[E]    centres whose keys outgrew the LDS histogram (and the next levels')
  DevBuf big2;     // int32[E]    the other list of the level-by-level passes
  DevBuf bighash;  // uint32[]    their histograms in global memory (3 * 2^HB words per workgroup)
  DevBuf gtab;     // uint32[2^(h+1)] class hyperedges by hub-mask subset sums, per light flag
  DevBuf slow;     // uint32[]    pair types with too many hubs for the subset sums
  // every buffer above, for release() and the device-memory accounting alike
  template <typename Self, typename F>
  static void each_of(Self& s, F f) {
    for (auto* b : {&s.hubidx, &s.hmask, &s.lsz, &s.ckey, &s.ckey2, &s.ukey, &s.ucnt, &s.nruns, &s.hmr, &s.lszr,
                    &s.types, &s.tcraw, &s.types2, &s.tcsrt, &s.tkey, &s.tcnt, &s.rkey, &s.first, &s.firstids,
                    &s.accd, &s.accr, &s.big, &s.big2, &s.bighash, &s.gtab, &s.slow})
      f(*b);
  }
  template <typename F>
  void each(F f) { each_of(*this, f); }
  template <typename F>
  void each(F f) const { each_of(*this, f); }
  void release() {
    each([](DevBuf& b) { b.release(); });
  }
  // the pair-type buffers, once (B) has run (they are the largest of a big graph's count)
  void release_types() {
    for (DevBuf* b : {&types, &tcraw, &types2, &tcsrt, &bighash}) b->release();
  }
};

// Scratch of the dynamic path (full projection rows, batch staging), reused across updates.
struct DynScratch {
  DevBuf deg, roff, rows, flag, ids, acc, tmp;
  DevBuf ins_off, ins_nodes, sizes, noff, nnodes, neoff, nsize, inc_node, inc_edge, srt_node,
      srt_edge, nvoff;
  DevBuf slen, soff, scur, spm;  // tiled delta: per (changed f, neighbour b) read position and mask
  DevBuf ord;                    // one-table delta: changed ids, most projected degree first
  // deg / roff / rows hold the full projection rows of the graph as it is now, built with tile
  // rows_tile (0 = one table). Cleared whenever the store changes.
  bool rows_valid = false;
  int32_t rows_tile = -1;
  unsigned long long rows_P = 0;  // entries of the held rows
  void release() {
    rows_valid = false;
    for (DevBuf* b : {&deg, &roff, &rows, &flag, &ids, &acc, &tmp, &ins_off, &ins_nodes, &sizes,
                      &noff, &nnodes, &neoff, &nsize, &inc_node, &inc_edge, &srt_node, &srt_edge,
                      &nvoff, &slen, &soff, &scur, &spm, &ord})
      b->release();
  }
};

// Device-side header of a captured update: the batch's sizes, written by the host before each
// launch, and what the launch reports back.
struct UpdHdr {
  int32_t E_old, E_new, n_del, n_ins;
  int32_t V_new, parts_del, parts_ins;
  int32_t overflow;           // set on the device: rows after the batch exceed the plan
  unsigned long long I_new;   // incidences after the batch
  unsigned long long P_new;   // projection-row entries after the batch
};

// Captured update (hm_graph_set_update_capture): capacities, the instantiated graph and the
// buffer addresses it was captured with (a graph launch is only valid while they are unchanged).
struct UpdPlan {
  bool on = false;
  bool disabled = false;      // the capture held a memset node: updates stay eager
  CapturedGraph cap;          // the instantiated update (cap.exec null: none)
  bool wide = false;          // 64-bit masks
  int32_t Ecap = 0, Vcap = 0, dcap = 0, icap = 0;
  int64_t Icap = 0, iicap = 0;
  unsigned long long Pcap = 0;
  int bits = 0;               // radix-sort key bits: every node id and the padding key fit
  unsigned grid_del = 0, grid_ins = 0;
  std::vector<const void*> sig;
  DevBuf in;                  // device input block: header, deleted ids, insert offsets, nodes
  PinnedBuf hin, hout;        // host staging of the input block; acc and header written back
  int32_t captures = 0;
  bool last_captured = false, last_overflow = false;
  void release() {
    cap.reset();
    in.release();
    hin.release();
    hout.release();
    sig.clear();
  }
};

}  // namespace hm

struct hm_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hm::DeviceCsr csr;
  hm::CountScratch scratch;
  hm::DynScratch dyn;
  // live[id] = 1 for hyperedges that have not been deleted
  std::vector<uint8_t> live;
  hm_graph_stats stats{};
  bool have_stats = false;
  // a hub-aware count leaves its projection statistics to the first hm_graph_get_stats (they are
  // not part of the bins): stats_lazy with the count's bins, for that call to complete `stats`
  bool stats_lazy = false;
  uint64_t lazy_bins[HM_STATIC_BINS] = {};
  float ms_project = 0.f, ms_triangles = 0.f;
  // id tile of the static count: 0 = automatic (one LDS table when E fits), > 0 = forced tiles
  int32_t tile = 0;
  int32_t last_tile = 0;  // tile used by the last count (0 = single table)
  // triples that held a deleted hyperedge before / hold an inserted one after the last update
  // with a delta (-1 when there was none)
  int64_t upd_removed = -1, upd_added = -1, upd_entries = -1;
  float ms_upd_rows = 0.f, ms_upd_delta = 0.f, ms_upd_store = 0.f;
  // captured static count (hm_graph_set_replay): replayed while the graph is unchanged
  uint64_t version = 0;  // bumped by every store update and tile change
  bool replay_on = false;
  hm::CapturedGraph replay;  // replay.exec null: no captured count
  uint64_t replay_version = 0;
  bool last_replayed = false;
  // one-table triangle form: -1 = autotuned per shape (tri_retune: tune again on the next count)
  int32_t tri_form = -1;
  bool tri_retune = false;
  int32_t last_tri_form = -1;  // form of the last one-table count (-1: none or tiled)
  bool last_tri_tuned = false;
  hm::UpdPlan uplan;  // captured update with a delta
  // hub-aware static count (hm_graph_set_hubs): 0 = off, -1 = automatic, h > 0 = h hubs
  int32_t hubs = 0;
  int32_t hub_auto_h = 0;          // hub count the automatic mode settled on (0: none yet)
  uint64_t hub_auto_version = 0;   // store version it was chosen for
  hm::HubScratch hub;
  hm_hub_stats hub_stats{};
  std::vector<int32_t> hub_nodes;  // hub set of the last hub-aware count (most frequent first)
};

namespace hm {
// Most hubs the hub-aware count takes (a neighbour key holds the hub pattern in 23 bits beside the
// overlap and a flag).
constexpr int32_t kHubMax = 23;
constexpr int32_t kHubDeltaHubs = 12;  // most hubs of an automatic hub-aware delta
// One run of the hub-aware pipeline: an optional pinned hub set, ids ranked first in the light
// order (with only_first, (D) runs over those centres alone: the light triangles holding one of
// them), and whether the projection statistics are computed.
struct HubRun {
  const std::vector<int32_t>* hub_nodes = nullptr;
  std::vector<int32_t>* hubs_out = nullptr;  // receives the hub set used
  const int32_t* first = nullptr;
  int32_t n_first = 0;
  bool only_first = false;
  bool stats = true;
  bool abc = true;        // false: (A), (B), (C) skipped (their bins stay 0; no key histograms)
  int64_t d_begin = 0;    // (D) over the centres of light ranks [d_begin, d_end) (d_end < 0: all)
  int64_t d_end = -1;
  double seconds = 0;     // > 0: (D)'s launches stop once this long has passed since the count began
  int64_t* d_done = nullptr;  // receives the end of the light ranks (D) enumerated
};
// Parts (A) + (B) + (C) into h_abc's bins (plus the projection statistics when run.stats) and
// (D) into h_d's bins (kAccLightTri: the light triangles enumerated).
hm_status hub_parts(hm_graph* g, int32_t h, const HubRun& run, unsigned long long* h_abc, unsigned long long* h_d,
                    bool* key_overflow);
// Hub-aware static count of g with h hubs (1..kHubMax) into the accumulator layout of the
// triangle path (30 bins; the projection statistics are not computed). *key_overflow is set when
// some hyperedge's light neighbours hold more distinct keys than its LDS histogram (fewer hubs
// then fit).
hm_status count_static_hub(hm_graph* g, int32_t h, unsigned long long* h_acc, bool* key_overflow,
                           bool stats = true);
// Hub count minimising the hub-aware cost model.
hm_status hub_choose(hm_graph* g, int32_t* h_out, double* cost_out);
// Projection statistics (sum and max of the projected degrees, wedges, oriented row sum, largest
// node degree) into their accumulator slots, without rows or triangles.
hm_status projection_stats(hm_graph* g, unsigned long long* h_acc);
}  // namespace hm
