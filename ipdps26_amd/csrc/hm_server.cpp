// Serving: a request loop over many device-resident hypergraphs.
//
// The reference has no serving layer: one host thread on the default stream with a synchronize
// after every launch (dynamicHyperGraph/structure/operations.cu:24-29, src/main.cu:238-250). This
// one keeps many graphs resident in HBM and runs their requests concurrently:
//   * admission — a graph is admitted only if its estimated device footprint (CSR, projection
//     rows bounded by sum_v deg_v (deg_v - 1), per-hyperedge scratch) fits the remaining HBM
//     budget; otherwise hm_server_add_graph returns HM_ERR_OVER_BUDGET;
//   * continuous batching — `slots` worker threads take whichever graph has pending work as soon
//     as they are free; every graph owns its HIP stream, so requests on different graphs overlap
//     on the device while requests on one graph keep their arrival order (an update is seen by
//     every later count);
//   * coalescing — consecutive count queries on an unchanged graph are answered by one count;
//   * kept counts (maintain_counts) — once a graph has been counted, each update adds its exact
//     delta to the graph's counts, and count queries are answered from them with no launch;
//   * batched counts (batch_graphs) — a worker whose next request is a count of a small graph
//     (last count under 1 ms) also takes other ready small graphs whose next request is a count
//     (up to batch_graphs graphs and about 4 ms of expected work) and counts them all with one
//     hm_batch launch sequence;
//   * captured updates (capture_updates) — each graph's update with its delta is one hipGraph
//     launch and one wait (hm_graph_set_update_capture);
//   * latency — each request records submit, start and finish times (steady clock).
// Counts are exact, so a result is the same alone and in the mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hm_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  int32_t graph = -1;
  int32_t kind = 0;  // HM_REQ_COUNT or HM_REQ_UPDATE
  std::vector<int32_t> dels;
  std::vector<int64_t> ins_off;
  std::vector<int32_t> ins_nodes;
  Clock::time_point submit, start, finish;
  hm_status status = HM_OK;
  std::string error;
  int64_t counts[HM_STATIC_BINS] = {0};
  int32_t first_new_id = -1, n_new = 0;
  bool done = false;
};

struct GraphSlot {
  hm_graph* g = nullptr;
  int64_t bytes = 0;
  std::deque<std::shared_ptr<Request>> pending;
  bool busy = false;     // a worker is running one of its requests
  bool queued = false;   // the graph id sits in the ready queue
  bool removed = false;
  double count_ms = -1.0;    // duration of the graph's last count (its share of a batched one)
  bool hub = false;          // counted and updated hub-aware (never batched)
  bool have_counts = false;  // maintain_counts: `counts` equal the graph's counts after its last
                             // finished request
  uint64_t counts[HM_STATIC_BINS] = {0};
};

// Device footprint of a graph: CSR, projection rows (4 B per directed pair, at most E^2) and
// per-hyperedge scratch. A hub-aware graph holds no full rows: its light upper rows (at most 12 B
// per entry, pairs through nodes other than the 8 most frequent) and pair-type entries (16 B, at
// most 1,024 per hyperedge before they are counted exactly) instead.
int64_t estimate_bytes(const int64_t* off, const int32_t* nodes, int32_t E, bool hub) {
  const int64_t I = off[E];
  std::unordered_map<int32_t, int64_t> deg;
  deg.reserve((size_t)std::min<int64_t>(I, 1 << 22));
  for (int64_t k = 0; k < I; ++k) deg[nodes[k]]++;
  std::vector<int64_t> d;
  d.reserve(deg.size());
  for (auto& kv : deg) d.push_back(kv.second);
  const size_t top = hub ? std::min<size_t>(8, d.size()) : 0;
  std::partial_sort(d.begin(), d.begin() + top, d.end(), std::greater<int64_t>());
  int64_t pairs = 0;
  for (size_t k = top; k < d.size(); ++k) pairs += d[k] * (d[k] - 1);
  const int64_t csr = 4 * (2 * I + 2 * (int64_t)E + 2 + 2 * (int64_t)deg.size());
  const int64_t rows = hub ? 6 * pairs + 16 * std::min<int64_t>(pairs / 2 + (1 << 20), 1024 * (int64_t)E)
                           : 4 * std::min<int64_t>(pairs, (int64_t)E * (int64_t)E);
  return csr + rows + 64 * (int64_t)E + (1 << 20);
}

}  // namespace

struct hm_server {
  hm_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  // a deque: workers and hm_server_remove_graph keep references to a slot while the lock is
  // released, and adding a graph must not move the others
  std::deque<GraphSlot> graphs;
  std::deque<int32_t> ready;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
  std::vector<std::thread> workers;
  uint64_t next_ticket = 1;
  int64_t used_bytes = 0;
  int64_t completed = 0, coalesced = 0, maintained = 0, batched_counts = 0, batch_launches = 0;
  bool stop = false;

  void enqueue_locked(int32_t gid) {
    GraphSlot& s = graphs[(size_t)gid];
    if (!s.busy && !s.queued && !s.pending.empty()) {
      s.queued = true;
      ready.push_back(gid);
      work_cv.notify_one();
    }
  }

  // Answers count requests from the slot's kept counts (s->mu held).
  void answer_kept_locked(GraphSlot& s, const std::vector<std::shared_ptr<Request>>& batch) {
    const auto now = Clock::now();
    for (auto& q : batch) {
      q->start = q->finish = now;
      q->status = HM_OK;
      for (int k = 0; k < HM_STATIC_BINS; ++k) q->counts[k] = (int64_t)s.counts[k];
      q->done = true;
    }
    completed += (int64_t)batch.size();
    maintained += (int64_t)batch.size();
    done_cv.notify_all();
  }

  // A count that may join a batch: one whose last count (alone, or its share of a batch) took
  // under kBatchSmallMs. Larger counts fill the device by themselves, and batching them would only
  // make every request of the batch wait for the whole batch.
  static constexpr double kBatchSmallMs = 1.0;
  static constexpr double kBatchBudgetMs = 4.0;  // expected duration of one batch at most
  static bool small_count(const GraphSlot& g) {
    return !g.hub && g.count_ms >= 0.0 && g.count_ms < kBatchSmallMs;
  }

  // Takes the head request of graph gid, plus every directly following count when the head is a
  // count (s->mu held; the slot becomes busy).
  std::vector<std::shared_ptr<Request>> take_head_locked(int32_t gid) {
    GraphSlot& s = graphs[(size_t)gid];
    s.queued = false;
    s.busy = true;
    std::vector<std::shared_ptr<Request>> batch{s.pending.front()};
    s.pending.pop_front();
    if (batch[0]->kind == HM_REQ_COUNT)
      while (!s.pending.empty() && s.pending.front()->kind == HM_REQ_COUNT) {
        batch.push_back(s.pending.front());
        s.pending.pop_front();
      }
    return batch;
  }

  // Records a finished request group of graph gid and releases the slot (s->mu held).
  void finish_locked(int32_t gid, std::vector<std::shared_ptr<Request>>& batch, hm_status st,
                     const std::string& err, Clock::time_point t0, Clock::time_point t1) {
    GraphSlot& s = graphs[(size_t)gid];
    Request& r = *batch[0];
    if (cfg.maintain_counts) {
      if (st != HM_OK) {
        s.have_counts = false;
      } else if (r.kind == HM_REQ_COUNT) {
        for (int k = 0; k < HM_STATIC_BINS; ++k) s.counts[k] = (uint64_t)r.counts[k];
        s.have_counts = true;
      } else if (s.have_counts) {
        for (int k = 0; k < HM_STATIC_BINS; ++k) s.counts[k] += (uint64_t)r.counts[k];
      }
    }
    for (auto& q : batch) {
      q->start = t0;
      q->finish = t1;
      q->status = st;
      q->error = err;
      if (q.get() != &r) std::memcpy(q->counts, r.counts, sizeof(r.counts));
      q->done = true;
    }
    completed += (int64_t)batch.size();
    coalesced += (int64_t)batch.size() - 1;
    s.busy = false;
    enqueue_locked(gid);
  }

  void worker_main() {
    (void)hipSetDevice(device);
    (void)hm_init(device);       // handles made by this thread live on the server's device
    hm_batch* counter = nullptr;  // batched counts of this worker (made on first use)
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      work_cv.wait(lk, [&] { return stop || !ready.empty(); });
      if (stop && ready.empty()) break;
      const int32_t gid = ready.front();
      ready.pop_front();
      GraphSlot& s = graphs[(size_t)gid];
      if (s.pending.empty()) {
        s.queued = false;
        continue;
      }
      std::vector<int32_t> gids{gid};
      std::vector<std::vector<std::shared_ptr<Request>>> groups{take_head_locked(gid)};
      if (groups[0][0]->kind == HM_REQ_COUNT && s.have_counts) {
        answer_kept_locked(s, groups[0]);
        s.busy = false;
        enqueue_locked(gid);
        continue;
      }
      // continuous batching: other ready graphs whose next request is a count join this count
      if (groups[0][0]->kind == HM_REQ_COUNT && cfg.batch_graphs > 1 && small_count(s)) {
        double expect_ms = s.count_ms;
        for (auto it = ready.begin(); it != ready.end() && (int32_t)gids.size() < cfg.batch_graphs;) {
          // (a graph with kept counts is left to its own turn: it needs no launch)
          GraphSlot& o = graphs[(size_t)*it];
          if (o.busy || o.have_counts || o.pending.empty() || o.pending.front()->kind != HM_REQ_COUNT ||
              !small_count(o) || expect_ms + o.count_ms > kBatchBudgetMs) {
            ++it;
            continue;
          }
          expect_ms += o.count_ms;
          gids.push_back(*it);
          groups.push_back(take_head_locked(*it));
          it = ready.erase(it);
        }
      }
      std::vector<hm_graph*> hs;
      for (int32_t g : gids) hs.push_back(graphs[(size_t)g].g);
      const bool over_budget = used_bytes > cfg.hbm_budget_bytes;
      lk.unlock();
      const auto t0 = Clock::now();
      Request& r = *groups[0][0];
      hm_status st;
      std::string err;
      if (r.kind == HM_REQ_COUNT && hs.size() > 1) {
        std::vector<uint64_t> out(hs.size() * HM_STATIC_BINS);
        st = counter ? HM_OK : hm_batch_create(&counter);
        if (st == HM_OK) st = hm_batch_count_static(counter, hs.data(), (int32_t)hs.size(), out.data());
        if (st != HM_OK) err = hm_last_error();
        for (size_t k = 0; k < hs.size(); ++k)
          for (int i = 0; i < HM_STATIC_BINS; ++i)
            groups[k][0]->counts[i] = st == HM_OK ? (int64_t)out[k * HM_STATIC_BINS + i] : 0;
      } else if (r.kind == HM_REQ_COUNT) {
        uint64_t out[HM_STATIC_BINS];
        st = hm_count_static(hs[0], out);
        for (int k = 0; k < HM_STATIC_BINS; ++k) r.counts[k] = (int64_t)out[k];
      } else if (over_budget) {
        // the graphs' buffers have grown past the budget: no update runs until a graph is removed
        st = HM_ERR_OVER_BUDGET;
        err = "hm_server: device memory held by the graphs exceeds the HBM budget; remove a graph";
      } else {
        std::vector<int32_t> ids(r.ins_off.size() - 1);
        st = hm_graph_update(hs[0], r.dels.data(), (int32_t)r.dels.size(), r.ins_off.data(),
                             r.ins_nodes.data(), (int32_t)ids.size(), ids.data(), r.counts);
        r.n_new = (int32_t)ids.size();
        r.first_new_id = ids.empty() ? -1 : ids[0];
      }
      if (st != HM_OK && err.empty()) err = hm_last_error();
      // what the graphs hold now (buffers grow with updates and first counts), charged below
      std::vector<int64_t> held;
      for (hm_graph* h : hs) held.push_back(hm::graph_device_bytes(h));
      const auto t1 = Clock::now();
      lk.lock();
      for (size_t k = 0; k < gids.size(); ++k) {
        GraphSlot& gs = graphs[(size_t)gids[k]];
        if (held[k] > gs.bytes) {
          used_bytes += held[k] - gs.bytes;
          gs.bytes = held[k];
        }
      }
      if (r.kind == HM_REQ_COUNT && st == HM_OK) {
        const double share = std::chrono::duration<double, std::milli>(t1 - t0).count() / (double)gids.size();
        for (int32_t g : gids) graphs[(size_t)g].count_ms = share;
      }
      for (size_t k = 0; k < gids.size(); ++k) finish_locked(gids[k], groups[k], st, err, t0, t1);
      if (gids.size() > 1) {
        batched_counts += (int64_t)gids.size();
        ++batch_launches;
      }
      done_cv.notify_all();
    }
    lk.unlock();
    if (counter) hm_batch_destroy(counter);
  }
};

extern "C" {

hm_status hm_server_create(const hm_server_config* cfg, hm_server** out) {
  hm::clear_error();
  if (!cfg || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_create: null argument");
  if (cfg->slots < 1 || cfg->slots > 64 || cfg->hbm_budget_bytes <= 0 ||
      (cfg->maintain_counts != 0 && cfg->maintain_counts != 1) || cfg->batch_graphs < 0 ||
      (cfg->capture_updates != 0 && cfg->capture_updates != 1) || cfg->hub_min_edges < 0)
    return hm::fail(HM_ERR_BAD_INPUT,
                    "hm_server_create: need 1 <= slots <= 64, a positive budget, maintain_counts 0 or 1, "
                    "batch_graphs >= 0, capture_updates 0 or 1 and hub_min_edges >= 0");
  *out = nullptr;
  size_t free_b = 0, total_b = 0;
  if (hipSetDevice(hm::current_device()) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
    return hm::fail(HM_ERR_DEVICE, "hm_server_create: no usable device");
  if ((uint64_t)cfg->hbm_budget_bytes > (uint64_t)total_b)
    return hm::fail(HM_ERR_OVER_BUDGET, "hm_server_create: budget exceeds device memory");
  hm_server* s = new (std::nothrow) hm_server();
  if (!s) return hm::fail(HM_ERR_OOM, "hm_server_create: host allocation failed");
  s->cfg = *cfg;
  s->device = hm::current_device();
  for (int32_t k = 0; k < cfg->slots; ++k) s->workers.emplace_back([s] { s->worker_main(); });
  *out = s;
  return HM_OK;
}

hm_status hm_server_destroy(hm_server* s) {
  if (!s) return HM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->work_cv.notify_all();
  }
  for (auto& t : s->workers) t.join();
  for (auto& g : s->graphs)
    if (g.g) hm_graph_destroy(g.g);
  delete s;
  return HM_OK;
}

hm_status hm_server_add_graph(hm_server* s, const int64_t* offsets, const int32_t* nodes,
                              int32_t num_edges, int32_t* graph_id) {
  hm::clear_error();
  if (!s || !graph_id || !offsets || num_edges < 0)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_server_add_graph: bad argument");
  const bool hub = s->cfg.hub_min_edges > 0 && num_edges >= s->cfg.hub_min_edges;
  const int64_t need = estimate_bytes(offsets, nodes, num_edges, hub);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->used_bytes + need > s->cfg.hbm_budget_bytes) {
      hm::set_error("hm_server_add_graph: needs ~%lld bytes, %lld of %lld in use",
                    (long long)need, (long long)s->used_bytes, (long long)s->cfg.hbm_budget_bytes);
      return HM_ERR_OVER_BUDGET;
    }
    s->used_bytes += need;  // reserved while the graph is built
  }
  hm_graph* g = nullptr;
  hm_status st = hm_graph_create(offsets, nodes, num_edges, 0, &g);
  if (st == HM_OK && s->cfg.capture_updates) st = hm_graph_set_update_capture(g, 1);
  if (st == HM_OK && hub) st = hm_graph_set_hubs(g, -1);
  if (st != HM_OK && g) {
    hm_graph_destroy(g);
    g = nullptr;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (st != HM_OK) {
    s->used_bytes -= need;
    return st;
  }
  GraphSlot slot;
  slot.g = g;
  slot.bytes = need;
  slot.hub = hub;
  s->graphs.push_back(std::move(slot));
  *graph_id = (int32_t)s->graphs.size() - 1;
  return HM_OK;
}

hm_status hm_server_remove_graph(hm_server* s, int32_t gid) {
  hm::clear_error();
  if (!s) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_remove_graph: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (gid < 0 || gid >= (int32_t)s->graphs.size() || s->graphs[(size_t)gid].removed)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_server_remove_graph: unknown graph");
  GraphSlot& g = s->graphs[(size_t)gid];
  s->done_cv.wait(lk, [&] { return !g.busy && g.pending.empty(); });
  hm_graph_destroy(g.g);
  g.g = nullptr;
  g.removed = true;
  s->used_bytes -= g.bytes;
  return HM_OK;
}

static hm_status submit(hm_server* s, std::shared_ptr<Request> r, uint64_t* ticket) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->graph < 0 || r->graph >= (int32_t)s->graphs.size() || s->graphs[(size_t)r->graph].removed)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_server_submit: unknown graph");
  r->ticket = s->next_ticket++;
  r->submit = Clock::now();
  s->requests[r->ticket] = r;
  GraphSlot& g = s->graphs[(size_t)r->graph];
  if (r->kind == HM_REQ_COUNT && g.have_counts && !g.busy && g.pending.empty()) {
    s->answer_kept_locked(g, {r});  // nothing ahead of it: the kept counts are current
  } else {
    g.pending.push_back(r);
    s->enqueue_locked(r->graph);
  }
  *ticket = r->ticket;
  return HM_OK;
}

hm_status hm_server_submit_count(hm_server* s, int32_t graph_id, uint64_t* ticket) {
  hm::clear_error();
  if (!s || !ticket) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_submit_count: null argument");
  auto r = std::make_shared<Request>();
  r->graph = graph_id;
  r->kind = HM_REQ_COUNT;
  return submit(s, r, ticket);
}

hm_status hm_server_submit_update(hm_server* s, int32_t graph_id, const int32_t* del_ids,
                                  int32_t n_del, const int64_t* ins_offsets,
                                  const int32_t* ins_nodes, int32_t n_ins, uint64_t* ticket) {
  hm::clear_error();
  if (!s || !ticket || n_del < 0 || n_ins < 0 || (n_del && !del_ids) || !ins_offsets)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_server_submit_update: bad argument");
  auto r = std::make_shared<Request>();
  r->graph = graph_id;
  r->kind = HM_REQ_UPDATE;
  r->dels.assign(del_ids, del_ids + n_del);
  r->ins_off.assign(ins_offsets, ins_offsets + n_ins + 1);
  const int64_t I = ins_offsets[n_ins];
  if (I < 0 || (I > 0 && !ins_nodes)) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_submit_update: bad CSR");
  r->ins_nodes.assign(ins_nodes, ins_nodes + I);
  return submit(s, r, ticket);
}

// Copies a finished request into out and forgets its ticket (s->mu held by the caller).
static hm_status take_result(hm_server* s, std::unordered_map<uint64_t, std::shared_ptr<Request>>::iterator it,
                             hm_result* out) {
  std::shared_ptr<Request> r = it->second;
  s->requests.erase(it);
  using ms = std::chrono::duration<double, std::milli>;
  out->status = r->status;
  out->kind = r->kind;
  std::memcpy(out->counts, r->counts, sizeof(out->counts));
  out->queued_ms = ms(r->start - r->submit).count();
  out->run_ms = ms(r->finish - r->start).count();
  out->latency_ms = ms(r->finish - r->submit).count();
  out->first_new_id = r->first_new_id;
  out->n_new = r->n_new;
  if (r->status != HM_OK) {
    hm::set_error("%s", r->error.c_str());
    return r->status;
  }
  return HM_OK;
}

hm_status hm_server_wait(hm_server* s, uint64_t ticket, hm_result* out) {
  hm::clear_error();
  if (!s || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_wait: unknown ticket");
  std::shared_ptr<Request> r = it->second;
  s->done_cv.wait(lk, [&] { return r->done; });
  return take_result(s, s->requests.find(ticket), out);
}

hm_status hm_server_poll(hm_server* s, uint64_t ticket, int32_t* done, hm_result* out) {
  hm::clear_error();
  if (!s || !done || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->requests.find(ticket);
  if (it == s->requests.end()) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_poll: unknown ticket");
  *done = it->second->done ? 1 : 0;
  if (!*done) return HM_OK;
  return take_result(s, it, out);
}

hm_status hm_server_get_stats(hm_server* s, hm_server_stats* out) {
  hm::clear_error();
  if (!s || !out) return hm::fail(HM_ERR_BAD_INPUT, "hm_server_get_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  out->graphs = 0;
  for (auto& g : s->graphs) out->graphs += g.removed ? 0 : 1;
  out->reserved_bytes = s->used_bytes;
  out->budget_bytes = s->cfg.hbm_budget_bytes;
  out->completed = s->completed;
  out->coalesced = s->coalesced;
  out->maintained = s->maintained;
  out->batched_counts = s->batched_counts;
  out->batch_launches = s->batch_launches;
  return HM_OK;
}

}  // extern "C"
