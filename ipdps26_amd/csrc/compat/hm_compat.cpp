// Source-compatible implementation of the reference's store and counting interfaces
// (dynamicHyperGraph/include/structure.hpp, motif.hpp, motif_update.hpp;
// temporalDynamicHMotif/include/temporal_structure.hpp, temporal_adjacency.hpp, temporal_count.hpp)
// over the libhmotif C ABI. Rows are kept host-side per key; every count builds device-resident
// graphs from the H2V rows and runs on the GPU through hm_count_static / hm_count_delta /
// hm_count_strict_inc. Nothing here counts on the CPU.
#include <climits>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>
#include <utility>

#include "hmotif.h"
#include "motif.hpp"
#include "motif_update.hpp"
#include "structure.hpp"
#include "temporal_count.hpp"

namespace {

struct RowStore {
  std::map<int, std::vector<int>> rows;  // key -> row (set semantics, insertion order kept)
};

RowStore* store_of(const CBSTContext& ctx) {
  if (!ctx.store) throw hmotif_compat_error("CBST context used before construct()");
  return static_cast<RowStore*>(ctx.store);
}

void sync_meta(CBSTContext& ctx) {
  RowStore* s = store_of(ctx);
  ctx.numRecords = s->rows.empty() ? 0 : s->rows.rbegin()->first;
  size_t total = 0;
  for (const auto& kv : s->rows) total += kv.second.size();
  ctx.initialPayloadSize = (int)total;
}

void add_value(std::vector<int>& row, int v) {
  for (int x : row)
    if (x == v) return;
  row.push_back(v);
}

void check(hm_status st, const char* what) {
  if (st != HM_OK)
    throw hmotif_compat_error(std::string(what) + ": " + hm_status_string(st) + ": " + hm_last_error());
}

// Iterates (key, values) of a (keys, payload, inclusive prefix) batch.
template <typename F>
void for_batch(const std::vector<int>& keys, const std::vector<int>& payload,
               const std::vector<int>& prefix, F&& f) {
  if (prefix.size() != keys.size()) throw hmotif_compat_error("batch: prefix and keys differ in length");
  int lo = 0;
  for (size_t i = 0; i < keys.size(); ++i) {
    const int hi = prefix[i];
    if (hi < lo || hi > (int)payload.size()) throw hmotif_compat_error("batch: bad prefix sizes");
    f(keys[i], payload.data() + lo, hi - lo);
    lo = hi;
  }
}

// CSR of the H2V rows 1..n of a context: row r - 1 holds key r.
void rows_to_csr(const CBSTContext& ctx, int n, std::vector<int64_t>& off, std::vector<int32_t>& nodes) {
  RowStore* s = store_of(ctx);
  off.assign((size_t)n + 1, 0);
  nodes.clear();
  for (int r = 1; r <= n; ++r) {
    auto it = s->rows.find(r);
    if (it != s->rows.end())
      for (int v : it->second) {
        if (v < 0) throw hmotif_compat_error("negative vertex id");
        nodes.push_back(v);
      }
    off[(size_t)r] = (int64_t)nodes.size();
  }
}

struct GraphHandle {
  hm_graph* g = nullptr;
  GraphHandle(const std::vector<int64_t>& off, const std::vector<int32_t>& nodes) {
    check(hm_graph_create(off.data(), nodes.empty() ? nullptr : nodes.data(),
                          (int32_t)(off.size() - 1), 0, &g),
          "hm_graph_create");
  }
  ~GraphHandle() { hm_graph_destroy(g); }
};

}  // namespace

// ---------------------------------------------------------------------------------------------
// structure.hpp
// ---------------------------------------------------------------------------------------------
void constructCBST(int* keys, int* startOffsets, int numRecords, int* flatPayload, int flatPayloadSize,
                   int payloadCapacity, const char* datasetName, CBSTContext& ctx) {
  if (numRecords < 0 || (numRecords > 0 && (!keys || !startOffsets)) || flatPayloadSize < 0)
    throw hmotif_compat_error("constructCBST: bad arguments");
  if (!ctx.store) ctx.store = new RowStore();
  RowStore* s = store_of(ctx);
  s->rows.clear();
  for (int r = 0; r < numRecords; ++r) {
    std::vector<int>& row = s->rows[keys[r]];
    row.clear();
    for (int p = startOffsets[r]; p >= 0 && p < flatPayloadSize; ++p) {
      const int v = flatPayload[p];
      if (v == 0 || v == INT_MIN) break;  // the padded-row terminators
      add_value(row, v);
    }
  }
  ctx.fixedSize = payloadCapacity;
  ctx.datasetName = datasetName;
  sync_meta(ctx);
}

void fillCBST(const std::vector<int>& insertKeys, const std::vector<int>& insertPayload,
              const std::vector<int>& insertPrefixSizes, CBSTContext& ctx) {
  RowStore* s = store_of(ctx);
  for_batch(insertKeys, insertPayload, insertPrefixSizes, [&](int key, const int* v, int n) {
    std::vector<int>& row = s->rows[key];
    for (int q = 0; q < n; ++q) add_value(row, v[q]);
  });
  sync_meta(ctx);
}

void unfillCBST(const std::vector<int>& keysToUnfill, const std::vector<int>& valuesToRemove,
                const std::vector<int>& removePrefixSizes, CBSTContext& ctx) {
  RowStore* s = store_of(ctx);
  for_batch(keysToUnfill, valuesToRemove, removePrefixSizes, [&](int key, const int* v, int n) {
    auto it = s->rows.find(key);
    if (it == s->rows.end()) return;
    for (int q = 0; q < n; ++q)
      for (size_t t = 0; t < it->second.size(); ++t)
        if (it->second[t] == v[q]) {
          it->second.erase(it->second.begin() + (long)t);
          break;
        }
  });
  sync_meta(ctx);
}

void deleteCBST(const std::vector<int>& deleteKeys, CBSTContext& ctx) {
  RowStore* s = store_of(ctx);
  for (int k : deleteKeys) s->rows.erase(k);
  sync_meta(ctx);
}

void insertCBST(const std::vector<int>& newKeys, const std::vector<int>& newPayload,
                const std::vector<int>& newPrefixSizes, CBSTContext& ctx) {
  RowStore* s = store_of(ctx);
  for_batch(newKeys, newPayload, newPrefixSizes, [&](int key, const int* v, int n) {
    std::vector<int>& row = s->rows[key];
    row.clear();
    for (int q = 0; q < n; ++q) add_value(row, v[q]);
  });
  sync_meta(ctx);
}

CBSTOperations::CBSTOperations(const char* datasetName, int payloadCapacity, int alignment)
    : name_(datasetName ? datasetName : "") {
  ctx_.fixedSize = payloadCapacity;
  ctx_.alignment = alignment;
  ctx_.store = new RowStore();
  ctx_.datasetName = name_.c_str();
}

CBSTOperations::~CBSTOperations() { delete static_cast<RowStore*>(ctx_.store); }

CBSTOperations::CBSTOperations(CBSTOperations&& o) noexcept
    : ctx_(o.ctx_), name_(std::move(o.name_)), constructed_(o.constructed_) {
  ctx_.datasetName = name_.c_str();
  o.ctx_.store = nullptr;
}

CBSTOperations& CBSTOperations::operator=(CBSTOperations&& o) noexcept {
  if (this != &o) {
    delete static_cast<RowStore*>(ctx_.store);
    ctx_ = o.ctx_;
    name_ = std::move(o.name_);
    ctx_.datasetName = name_.c_str();
    constructed_ = o.constructed_;
    o.ctx_.store = nullptr;
  }
  return *this;
}

void CBSTOperations::construct(int* keys, int* startOffsets, int numRecords, int* flatPayload,
                               int flatPayloadSize) {
  constructCBST(keys, startOffsets, numRecords, flatPayload, flatPayloadSize, ctx_.fixedSize,
                name_.c_str(), ctx_);
  constructed_ = true;
}

void CBSTOperations::insert(const std::vector<int>& insertKeys, const std::vector<int>& insertPayload,
                            const std::vector<int>& insertPrefixSizes) {
  insertCBST(insertKeys, insertPayload, insertPrefixSizes, ctx_);
}

void CBSTOperations::erase(const std::vector<int>& deleteKeys) { deleteCBST(deleteKeys, ctx_); }

void CBSTOperations::findAndPrint(const std::vector<int>& ids) const {
  for (int id : ids) {
    std::cout << name_ << " " << id << ":";
    for (int v : cbstRow(ctx_, id)) std::cout << " " << v;
    std::cout << "\n";
  }
}

const CBSTContext& CBSTOperations::context() const { return ctx_; }

std::vector<int> cbstRow(const CBSTContext& ctx, int key) {
  RowStore* s = store_of(ctx);
  auto it = s->rows.find(key);
  return it == s->rows.end() ? std::vector<int>() : it->second;
}

// ---------------------------------------------------------------------------------------------
// motif.hpp / motif_update.hpp
// ---------------------------------------------------------------------------------------------
void computeMotifCounts(const CBSTContext& h2vCtx, const CBSTContext& v2hCtx, const CBSTContext& h2hCtx,
                        int numHyperedges, std::vector<unsigned long long>& outCounts30) {
  (void)v2hCtx;
  (void)h2hCtx;
  if (numHyperedges < 0) throw hmotif_compat_error("computeMotifCounts: negative hyperedge count");
  std::vector<int64_t> off;
  std::vector<int32_t> nodes;
  rows_to_csr(h2vCtx, numHyperedges, off, nodes);
  GraphHandle g(off, nodes);
  uint64_t out[HM_STATIC_BINS];
  check(hm_count_static(g.g, out), "hm_count_static");
  outCounts30.assign(out, out + HM_STATIC_BINS);
}

void computeMotifCounts(const CBSTContext& h2vCtx, const CBSTContext& v2hCtx, const CBSTContext& h2hCtx,
                        int numHyperedges) {
  std::vector<unsigned long long> c;
  computeMotifCounts(h2vCtx, v2hCtx, h2hCtx, numHyperedges, c);
  std::cout << "Motif counts (30 bins): ";
  for (int i = 0; i < HM_STATIC_BINS; ++i) std::cout << c[(size_t)i] << (i + 1 < HM_STATIC_BINS ? ' ' : '\n');
}

void computeMotifCountsDelta(const CBSTContext& oldH2vCtx, const CBSTContext& oldH2hCtx,
                             const CBSTContext& newH2vCtx, const CBSTContext& newH2hCtx,
                             const std::vector<int>& deletedHyperedgeIds,
                             const std::vector<int>& insertedHyperedgeIds,
                             std::vector<long long>& outDeltaCounts) {
  (void)oldH2hCtx;
  (void)newH2hCtx;
  std::vector<int64_t> off0, off1;
  std::vector<int32_t> nodes0, nodes1;
  rows_to_csr(oldH2vCtx, oldH2vCtx.numRecords, off0, nodes0);
  rows_to_csr(newH2vCtx, newH2vCtx.numRecords, off1, nodes1);
  std::vector<int32_t> del, ins;
  for (int id : deletedHyperedgeIds) del.push_back(id - 1);
  for (int id : insertedHyperedgeIds) ins.push_back(id - 1);
  GraphHandle g0(off0, nodes0), g1(off1, nodes1);
  int64_t out[HM_STATIC_BINS];
  check(hm_count_delta(g0.g, g1.g, del.empty() ? nullptr : del.data(), (int32_t)del.size(),
                       ins.empty() ? nullptr : ins.data(), (int32_t)ins.size(), out),
        "hm_count_delta");
  outDeltaCounts.assign(out, out + HM_STATIC_BINS);
}

void computeMotifCountsDelta(const CBSTContext& oldH2vCtx, const CBSTContext& oldH2hCtx,
                             const CBSTContext& newH2vCtx, const CBSTContext& newH2hCtx,
                             const std::vector<int>& deletedHyperedgeIds,
                             const std::vector<int>& insertedHyperedgeIds,
                             std::vector<int>& outDeltaCounts) {
  std::vector<long long> d;
  computeMotifCountsDelta(oldH2vCtx, oldH2hCtx, newH2vCtx, newH2hCtx, deletedHyperedgeIds,
                          insertedHyperedgeIds, d);
  for (size_t k = 0; k < d.size(); ++k)
    if (d[k] < (long long)INT_MIN || d[k] > (long long)INT_MAX)
      throw hmotif_compat_error("computeMotifCountsDelta: delta of bin " + std::to_string(k) + " (" +
                                std::to_string(d[k]) +
                                ") does not fit an int; use the std::vector<long long> overload");
  outDeltaCounts.assign(d.begin(), d.end());
}

// ---------------------------------------------------------------------------------------------
// temporal_structure.hpp / temporal_adjacency.hpp / temporal_count.hpp
// ---------------------------------------------------------------------------------------------
TemporalCBSTOperations::TemporalCBSTOperations(const std::string& label, int payloadCapacity, int alignment)
    : label_(label) {
  const char* names[3] = {"-Older", "-Middle", "-Newest"};
  for (int s = 0; s < 3; ++s)
    slots_[s].reset(new CBSTOperations((label_ + names[s]).c_str(), payloadCapacity, alignment));
}

TemporalCBSTOperations::~TemporalCBSTOperations() = default;

void TemporalCBSTOperations::rotate() { oldestIdx_ = (oldestIdx_ + 1) % 3; }

int TemporalCBSTOperations::resolveIndex(TemporalLayer layer) const {
  switch (layer) {
    case TemporalLayer::Older: return oldestIdx_;
    case TemporalLayer::Middle: return (oldestIdx_ + 1) % 3;
    case TemporalLayer::Newest: return (oldestIdx_ + 2) % 3;
  }
  return oldestIdx_;
}

void TemporalCBSTOperations::resetLayer(TemporalLayer layer) { context(layer).initialPayloadSize = 0; }

void TemporalCBSTOperations::construct(TemporalLayer layer, int* cbstKeys, int* cbstStartOffsets,
                                       int numRecords, int* flatPayload, int flatPayloadSize) {
  slots_[resolveIndex(layer)]->construct(cbstKeys, cbstStartOffsets, numRecords, flatPayload, flatPayloadSize);
}

void TemporalCBSTOperations::insert(TemporalLayer layer, const std::vector<int>& keys,
                                    const std::vector<int>& payload, const std::vector<int>& prefix) {
  slots_[resolveIndex(layer)]->insert(keys, payload, prefix);
}

void TemporalCBSTOperations::erase(TemporalLayer layer, const std::vector<int>& keys) {
  slots_[resolveIndex(layer)]->erase(keys);
}

void TemporalCBSTOperations::fill(TemporalLayer layer, const std::vector<int>& keys,
                                  const std::vector<int>& payload, const std::vector<int>& prefix) {
  fillCBST(keys, payload, prefix, context(layer));
}

void TemporalCBSTOperations::unfill(TemporalLayer layer, const std::vector<int>& keys,
                                    const std::vector<int>& valuesToRemove,
                                    const std::vector<int>& removePrefix) {
  unfillCBST(keys, valuesToRemove, removePrefix, context(layer));
}

const CBSTContext& TemporalCBSTOperations::context(TemporalLayer layer) const {
  return slots_[resolveIndex(layer)]->context();
}

CBSTContext& TemporalCBSTOperations::context(TemporalLayer layer) {
  return const_cast<CBSTContext&>(slots_[resolveIndex(layer)]->context());
}

TemporalLayer TemporalCBSTOperations::oldest() const { return static_cast<TemporalLayer>(oldestIdx_); }
TemporalLayer TemporalCBSTOperations::middle() const { return static_cast<TemporalLayer>((oldestIdx_ + 1) % 3); }
TemporalLayer TemporalCBSTOperations::newest() const { return static_cast<TemporalLayer>((oldestIdx_ + 2) % 3); }

CSRNeighbors buildCrossAdjacency(const CrossAdjacencySpec& spec, const TemporalCBSTOperations& h2v,
                                 const TemporalCBSTOperations& v2h) {
  (void)v2h;
  const CBSTContext& a = h2v.context(spec.anchorLayer);
  const CBSTContext& b = h2v.context(spec.neighborLayer);
  std::unordered_map<int, std::vector<int>> vert;  // vertex -> neighbour-layer hyperedges
  for (const auto& kv : store_of(b)->rows)
    for (int v : kv.second) vert[v].push_back(kv.first);
  CSRNeighbors out;
  out.numRows = (int)spec.anchorHyperedgeIds.size();
  out.rowPtr.assign(1, 0);
  for (int id : spec.anchorHyperedgeIds) {
    std::set<int> nb;
    for (int v : cbstRow(a, id)) {
      auto it = vert.find(v);
      if (it != vert.end()) nb.insert(it->second.begin(), it->second.end());
    }
    out.colIdx.insert(out.colIdx.end(), nb.begin(), nb.end());
    out.rowPtr.push_back((int)out.colIdx.size());
  }
  return out;
}

CSRNeighbors buildAdjOlderToMiddle(const std::vector<int>& anchorsMiddle, const TemporalHypergraphIndex& thg) {
  return buildCrossAdjacency({TemporalLayer::Middle, TemporalLayer::Older, anchorsMiddle}, thg.h2v, thg.v2h);
}

CSRNeighbors buildAdjMiddleToNewest(const std::vector<int>& anchorsMiddle, const TemporalHypergraphIndex& thg) {
  return buildCrossAdjacency({TemporalLayer::Middle, TemporalLayer::Newest, anchorsMiddle}, thg.h2v, thg.v2h);
}

CSRNeighbors buildAdjOlderToNewest(const std::vector<int>& anchorsOlder, const TemporalHypergraphIndex& thg) {
  return buildCrossAdjacency({TemporalLayer::Older, TemporalLayer::Newest, anchorsOlder}, thg.h2v, thg.v2h);
}

void computeTemporalMotifCountsStrictInc(const TemporalHypergraphIndex& thg, std::vector<int>& outCounts30) {
  std::vector<int64_t> off[3];
  std::vector<int32_t> nodes[3];
  const TemporalLayer layers[3] = {TemporalLayer::Older, TemporalLayer::Middle, TemporalLayer::Newest};
  const int64_t* po[3];
  const int32_t* pn[3];
  int32_t nr[3];
  for (int l = 0; l < 3; ++l) {
    const CBSTContext& c = thg.h2v.context(layers[l]);
    rows_to_csr(c, c.numRecords, off[l], nodes[l]);
    po[l] = off[l].data();
    pn[l] = nodes[l].empty() ? nullptr : nodes[l].data();
    nr[l] = (int32_t)(off[l].size() - 1);
  }
  uint64_t out[HM_STATIC_BINS];
  check(hm_count_strict_inc(po, pn, nr, out), "hm_count_strict_inc");
  for (int k = 0; k < HM_STATIC_BINS; ++k)
    if (out[k] > (uint64_t)INT_MAX)
      throw hmotif_compat_error("computeTemporalMotifCountsStrictInc: count of bin " + std::to_string(k) +
                                " (" + std::to_string(out[k]) + ") does not fit an int");
  outCounts30.assign(out, out + HM_STATIC_BINS);
}

void computeTemporalMotifCountsStrictIncDelta(const TemporalHypergraphIndex& oldWindow,
                                              const TemporalHypergraphIndex& newWindow,
                                              std::vector<int>& outDeltaCounts30) {
  std::vector<int> a, b;
  computeTemporalMotifCountsStrictInc(oldWindow, a);
  computeTemporalMotifCountsStrictInc(newWindow, b);
  outDeltaCounts30.assign(HM_STATIC_BINS, 0);
  for (int i = 0; i < HM_STATIC_BINS; ++i) outDeltaCounts30[(size_t)i] = b[(size_t)i] - a[(size_t)i];
}

bool applyTemporalDeltasFromFile(const std::string& path, TemporalHypergraphIndex& thg) {
  if (path.empty()) return false;
  std::ifstream in(path);
  if (!in) return false;
  std::map<int, std::vector<int>> add, rem;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    char op;
    int h;
    if (!(iss >> op >> h)) continue;
    std::vector<int> vs;
    for (int v; iss >> v;) vs.push_back(v);
    if (op == 'A') add[h].insert(add[h].end(), vs.begin(), vs.end());
    else if (op == 'R') rem[h].insert(rem[h].end(), vs.begin(), vs.end());
  }
  auto batch = [](const std::map<int, std::vector<int>>& m, std::vector<int>& k, std::vector<int>& p,
                  std::vector<int>& pre) {
    for (const auto& kv : m) {
      k.push_back(kv.first);
      p.insert(p.end(), kv.second.begin(), kv.second.end());
      pre.push_back((int)p.size());
    }
  };
  std::map<int, std::vector<int>> iadd, irem;  // mirrored: vertex -> hyperedges
  for (const auto& kv : add)
    for (int v : kv.second) iadd[v].push_back(kv.first);
  for (const auto& kv : rem)
    for (int v : kv.second) irem[v].push_back(kv.first);
  std::vector<int> k, p, pre;
  batch(rem, k, p, pre);
  if (!k.empty()) thg.h2v.unfill(TemporalLayer::Newest, k, p, pre);
  k.clear(); p.clear(); pre.clear();
  batch(add, k, p, pre);
  if (!k.empty()) thg.h2v.fill(TemporalLayer::Newest, k, p, pre);
  k.clear(); p.clear(); pre.clear();
  batch(irem, k, p, pre);
  if (!k.empty()) thg.v2h.unfill(TemporalLayer::Newest, k, p, pre);
  k.clear(); p.clear(); pre.clear();
  batch(iadd, k, p, pre);
  if (!k.empty()) thg.v2h.fill(TemporalLayer::Newest, k, p, pre);
  return true;
}
