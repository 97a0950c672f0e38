// hmotif.hpp — C++ host interface over the C ABI of libhmotif (include/hmotif.h).
//
// RAII handles plus the reference's entry-point names for the counting path:
//   computeMotifCounts        <- dynamicHyperGraph/include/motif.hpp:10-13 (30 bins, here uint64
//                                and returned instead of printed)
//   computeMotifCountsDelta   <- dynamicHyperGraph/include/motif_update.hpp:13-19 (old/new snapshots,
//                                deleted and inserted ids; here 0-based ids and int64 deltas)
//   computeTemporalMotifCounts <- THyMe+'s 96-bin count (THyMe-main/code/main_thymeP.cpp:219-367),
//                                the count temporalDynamicHMotif/include/temporal_count.hpp:10-11
//                                stands in for
// Errors are reported as hmotif::Error (status + message), never by exiting the process.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hmotif.h"

namespace hmotif {

class Error : public std::runtime_error {
 public:
  Error(hm_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
  hm_status status;
};

inline void check(hm_status st) {
  if (st != HM_OK)
    throw Error(st, std::string(hm_status_string(st)) + ": " + hm_last_error());
}

using StaticCounts = std::array<uint64_t, HM_STATIC_BINS>;
using StaticDelta = std::array<int64_t, HM_STATIC_BINS>;
using TemporalCounts = std::array<uint64_t, HM_TEMPORAL_BINS>;

// Hyperedge lists as a CSR: nodes of hyperedge e are nodes[offsets[e] .. offsets[e+1]).
struct Csr {
  std::vector<int64_t> offsets{0};
  std::vector<int32_t> nodes;
  int32_t size() const { return (int32_t)offsets.size() - 1; }
  void add(const std::vector<int32_t>& edge) {
    nodes.insert(nodes.end(), edge.begin(), edge.end());
    offsets.push_back((int64_t)nodes.size());
  }
};

class Graph {
 public:
  explicit Graph(const Csr& csr, int32_t num_nodes = 0) {
    check(hm_graph_create(csr.offsets.data(), csr.nodes.data(), csr.size(), num_nodes, &g_));
  }
  ~Graph() { hm_graph_destroy(g_); }
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  Graph(Graph&& o) noexcept : g_(std::exchange(o.g_, nullptr)) {}
  Graph& operator=(Graph&& o) noexcept {
    std::swap(g_, o.g_);
    return *this;
  }
  StaticCounts count() {
    StaticCounts out{};
    check(hm_count_static(g_, out.data()));
    return out;
  }
  // Deletes / inserts one batch; returns the ids given to the inserted hyperedges and fills
  // *delta (if not null) with counts(after) - counts(before).
  std::vector<int32_t> update(const std::vector<int32_t>& deleted, const Csr& inserted,
                              StaticDelta* delta = nullptr) {
    std::vector<int32_t> ids((size_t)inserted.size());
    check(hm_graph_update(g_, deleted.data(), (int32_t)deleted.size(), inserted.offsets.data(),
                          inserted.nodes.data(), inserted.size(), ids.data(),
                          delta ? delta->data() : nullptr));
    return ids;
  }
  hm_graph_stats stats() const {
    hm_graph_stats s{};
    check(hm_graph_get_stats(g_, &s));
    return s;
  }
  hm_graph* handle() const { return g_; }

 private:
  hm_graph* g_ = nullptr;
};

class TemporalGraph {
 public:
  TemporalGraph(const Csr& csr, const std::vector<int64_t>& times) {
    if ((int64_t)times.size() != csr.size()) throw Error(HM_ERR_BAD_INPUT, "one timestamp per hyperedge");
    check(hm_tgraph_create(csr.offsets.data(), csr.nodes.data(), times.data(), csr.size(), &g_));
  }
  ~TemporalGraph() { hm_tgraph_destroy(g_); }
  TemporalGraph(const TemporalGraph&) = delete;
  TemporalGraph& operator=(const TemporalGraph&) = delete;
  TemporalCounts count(int64_t delta) {
    TemporalCounts out{};
    check(hm_count_temporal(g_, delta, out.data()));
    return out;
  }
  hm_tgraph* handle() const { return g_; }

 private:
  hm_tgraph* g_ = nullptr;
};

// Reference entry-point names.
inline StaticCounts computeMotifCounts(Graph& g) { return g.count(); }

inline StaticDelta computeMotifCountsDelta(Graph& oldGraph, Graph& newGraph,
                                           const std::vector<int32_t>& deletedHyperedgeIds,
                                           const std::vector<int32_t>& insertedHyperedgeIds) {
  StaticDelta out{};
  check(hm_count_delta(oldGraph.handle(), newGraph.handle(), deletedHyperedgeIds.data(),
                       (int32_t)deletedHyperedgeIds.size(), insertedHyperedgeIds.data(),
                       (int32_t)insertedHyperedgeIds.size(), out.data()));
  return out;
}

inline TemporalCounts computeTemporalMotifCounts(TemporalGraph& g, int64_t delta) {
  return g.count(delta);
}

}  // namespace hmotif
