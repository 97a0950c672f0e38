// hm_ref_check — a driver written against the reference's own headers (structure.hpp, motif.hpp,
// motif_update.hpp, temporal_*.hpp, resolved to include/compat), following the flow of
// dynamicHyperGraph/src/main.cu:315-360 and temporalDynamicHMotif/src/main.cu:260-380:
//   1. H2V / V2H stores built from padded flat rows with 1-based ids, then computeMotifCounts;
//   2. one update batch (deletes, a reused id, appended ids) applied to a second set of stores,
//      then computeMotifCountsDelta, plus a recount of the updated stores;
//   3. the reference driver's synthetic temporal window (--temporal-synthetic), strict-inc count
//      and window delta;
//   4. applyTemporalDeltasFromFile on that window.
// Usage: hm_ref_check <mochy.csv> <delta_file>   (0-based vertex ids in the CSV, as MoCHy writes)
// Prints one line per result: "<label> (30 bins): c0 ... c29".
#include <climits>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "motif.hpp"
#include "motif_update.hpp"
#include "structure.hpp"
#include "temporal_count.hpp"

namespace {

using Rows = std::vector<std::vector<int>>;

// Flat payload with every row padded to a multiple of 4 by zeros and a closing INT_MIN, and the
// row start offsets (as the reference's flatten2DVector / prepareCBSTData produce them).
void flatten(const Rows& rows, std::vector<int>& flat, std::vector<int>& starts, std::vector<int>& keys) {
  flat.clear();
  starts.clear();
  keys.clear();
  for (size_t r = 0; r < rows.size(); ++r) {
    starts.push_back((int)flat.size());
    keys.push_back((int)r + 1);
    for (int v : rows[r]) flat.push_back(v);
    const size_t padded = ((rows[r].size() + 4) / 4) * 4;
    while (flat.size() - (size_t)starts.back() < padded) flat.push_back(0);
  }
  flat.push_back(INT_MIN);
}

Rows invert(const Rows& h2v) {
  Rows v2h;
  for (size_t e = 0; e < h2v.size(); ++e)
    for (int v : h2v[e]) {
      if ((size_t)v >= v2h.size()) v2h.resize((size_t)v + 1);
      v2h[(size_t)v].push_back((int)e + 1);
    }
  if (!v2h.empty()) v2h.erase(v2h.begin());  // vertex ids are 1-based: row 0 is unused
  return v2h;
}

void construct(CBSTOperations& ops, const Rows& rows) {
  std::vector<int> flat, starts, keys;
  flatten(rows, flat, starts, keys);
  ops.construct(keys.data(), starts.data(), (int)rows.size(), flat.data(), (int)flat.size());
}

template <typename T>
void print(const char* label, const std::vector<T>& c) {
  std::cout << label << " (30 bins):";
  for (const T& x : c) std::cout << ' ' << x;
  std::cout << '\n';
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: hm_ref_check <mochy.csv> <delta_file>\n";
    return 2;
  }
  try {
    Rows h2v;
    {
      std::ifstream in(argv[1]);
      for (std::string line; std::getline(in, line);) {
        if (line.empty()) continue;
        std::vector<int> row;
        std::stringstream ss(line);
        for (std::string tok; std::getline(ss, tok, ',');) row.push_back(std::stoi(tok) + 1);
        h2v.push_back(row);
      }
    }
    const int E = (int)h2v.size();
    // 1. static count
    CBSTOperations h2vOps("H2V", 64), v2hOps("V2H", 64), h2hOps("H2H", 64);
    construct(h2vOps, h2v);
    construct(v2hOps, invert(h2v));
    construct(h2hOps, Rows());
    std::vector<unsigned long long> before;
    computeMotifCounts(h2vOps.context(), v2hOps.context(), h2hOps.context(), E, before);
    print("Motif counts", before);

    // 2. one batch: delete ids 2, 5, 7; re-insert id 5 with the row of id 1; append two rows
    Rows upd = h2v;
    std::vector<int> deleted, inserted;
    for (int id : {2, 5, 7})
      if (id <= E) {
        deleted.push_back(id);
        upd[(size_t)id - 1].clear();
      }
    if (E >= 5) {
      upd[4] = h2v[0];
      inserted.push_back(5);
    }
    if (E >= 3) {
      upd.push_back(h2v[2]);
      inserted.push_back((int)upd.size());
      std::vector<int> mix = h2v[0];
      mix.push_back(h2v[1].empty() ? 1 : h2v[1][0]);
      upd.push_back(mix);
      inserted.push_back((int)upd.size());
    }
    CBSTOperations h2vNew("H2V-new", 64), v2hNew("V2H-new", 64), h2hNew("H2H-new", 64);
    construct(h2vNew, upd);
    construct(v2hNew, invert(upd));
    construct(h2hNew, Rows());
    std::vector<int> delta;
    computeMotifCountsDelta(h2vOps.context(), h2hOps.context(), h2vNew.context(), h2hNew.context(),
                            deleted, inserted, delta);
    print("Delta counts", delta);
    std::vector<unsigned long long> after;
    computeMotifCounts(h2vNew.context(), v2hNew.context(), h2hNew.context(), (int)upd.size(), after);
    print("Recount after", after);

    // 3. the reference driver's synthetic temporal window
    const Rows L = {{1, 2}, {2, 3}, {1, 3}}, L2 = {{1, 2}, {2, 3}, {1, 2, 3}};
    TemporalHypergraphIndex oldWin(64), newWin(64);
    std::vector<int> flat, starts, keys, flat2, starts2, keys2;
    flatten(L, flat, starts, keys);
    flatten(L2, flat2, starts2, keys2);
    for (TemporalLayer t : {TemporalLayer::Older, TemporalLayer::Middle, TemporalLayer::Newest}) {
      oldWin.h2v.construct(t, keys.data(), starts.data(), 3, flat.data(), (int)flat.size());
      const bool newest = t == TemporalLayer::Newest;
      newWin.h2v.construct(t, newest ? keys2.data() : keys.data(), newest ? starts2.data() : starts.data(), 3,
                           newest ? flat2.data() : flat.data(), (int)(newest ? flat2.size() : flat.size()));
    }
    std::vector<int> tc, td;
    computeTemporalMotifCountsStrictInc(oldWin, tc);
    print("Temporal strict-inc counts", tc);
    computeTemporalMotifCountsStrictIncDelta(oldWin, newWin, td);
    print("Temporal strict-inc delta", td);

    // 4. deltas from a file onto the new window's newest layer
    if (!applyTemporalDeltasFromFile(argv[2], newWin)) {
      std::cerr << "cannot read " << argv[2] << "\n";
      return 1;
    }
    std::vector<int> tf;
    computeTemporalMotifCountsStrictInc(newWin, tf);
    print("Temporal strict-inc after file deltas", tf);
    std::cout << "Newest row 3 after file deltas:";
    for (int v : cbstRow(newWin.h2v.context(TemporalLayer::Newest), 3)) std::cout << ' ' << v;
    std::cout << '\n';
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
