// Readers for the reference programs' input files (host code of the command-line tools).
//   MoCHy CSV (MoCHy-master/read_data.cpp:8-45): one hyperedge per line, comma-separated node
//     ids; a node repeated inside a line is kept once; lines are hyperedges in file order.
//   THyMe TSV (THyMe-main/code/read_data.cpp:22-58): "n1,n2,...<TAB>timestamp"; timestamps must
//     be integral (THyMe reads them as double; the engine keeps int64).
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/hmotif.hpp"

namespace hmio {

inline std::vector<int32_t> parse_ids(const std::string& s, const std::string& where) {
  std::vector<int32_t> out;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    size_t used = 0;
    long v = std::stol(tok, &used);
    if (v < 0 || v > INT32_MAX) throw std::runtime_error("node id out of range at " + where);
    bool seen = false;
    for (int32_t x : out) seen |= (x == (int32_t)v);
    if (!seen) out.push_back((int32_t)v);
  }
  return out;
}

inline hmotif::Csr read_mochy_csv(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  hmotif::Csr csr;
  std::string line;
  size_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    csr.add(parse_ids(line, path + ":" + std::to_string(lineno)));
  }
  return csr;
}

inline hmotif::Csr read_thyme_tsv(const std::string& path, std::vector<int64_t>* times) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  hmotif::Csr csr;
  std::string line;
  size_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    const size_t tab = line.find('\t');
    if (tab == std::string::npos) throw std::runtime_error("missing timestamp at line " + std::to_string(lineno));
    const std::string where = path + ":" + std::to_string(lineno);
    csr.add(parse_ids(line.substr(0, tab), where));
    const double t = std::stod(line.substr(tab + 1));
    if (t != std::floor(t) || std::fabs(t) > 4.0e18) throw std::runtime_error("non-integral timestamp at " + where);
    times->push_back((int64_t)t);
  }
  return csr;
}

}  // namespace hmio
