// Seeded synthetic hypergraph generator (host side of libhmotif).
//
// Replaces the reference's unseeded generator
// (dynamicHyperGraph/src/graphGeneration.cpp:10-51, srand(time(0)), sizes uniform on [1, m])
// with a reproducible one:
//   * hyperedge sizes follow a truncated discrete power law P(k) ∝ k^-alpha on [kmin, kmax];
//   * members are k distinct nodes drawn by successive sampling from a Zipf(s) popularity over
//     ranks 1..V (a repeated draw is redrawn); node id = rank - 1;
//   * sizes and members come from two independent xoshiro256** streams, so the number of
//     incidences is known from the size stream alone.
// Timestamps for temporal workloads are strictly increasing int64 microseconds.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hm_internal.hpp"
#include "hm_rng.hpp"

namespace {

// Stream seeds: size stream and member stream are derived from the user seed.
inline uint64_t size_seed(uint64_t seed) { return seed * 2ull + 0x51ull; }
inline uint64_t member_seed(uint64_t seed) { return seed * 2ull + 0xA7ull; }

hm_status check_params(const hm_gen_params* p) {
  if (!p) return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: null params");
  if (p->num_nodes <= 0 || p->num_nodes > (int64_t)INT32_MAX)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: num_nodes must be in [1, 2^31)");
  if (p->num_edges < 0 || p->num_edges > (int64_t)INT32_MAX)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: num_edges must be in [0, 2^31)");
  if (p->kmin < 1 || p->kmax < p->kmin)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: need 1 <= kmin <= kmax");
  if ((int64_t)p->kmax > p->num_nodes)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: kmax exceeds num_nodes");
  if (!(p->alpha >= 0.0) || !(p->zipf_s >= 0.0))
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen: alpha and zipf_s must be >= 0");
  return HM_OK;
}

// Draws the size sequence. Returns total incidences.
int64_t draw_sizes(const hm_gen_params* p, std::vector<int32_t>* sizes) {
  const int nk = p->kmax - p->kmin + 1;
  std::vector<double> cdf(nk);
  double acc = 0.0;
  for (int i = 0; i < nk; ++i) {
    acc += std::pow((double)(p->kmin + i), -p->alpha);
    cdf[i] = acc;
  }
  for (int i = 0; i < nk; ++i) cdf[i] /= acc;
  cdf[nk - 1] = 1.0;
  hm::Xoshiro256ss rng(size_seed(p->seed));
  int64_t total = 0;
  if (sizes) sizes->resize((size_t)p->num_edges);
  for (int64_t e = 0; e < p->num_edges; ++e) {
    double u = rng.uniform();
    int k = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    if (k >= nk) k = nk - 1;
    int32_t sz = p->kmin + k;
    if (sizes) (*sizes)[(size_t)e] = sz;
    total += sz;
  }
  return total;
}

}  // namespace

extern "C" {

hm_status hm_gen_static_size(const hm_gen_params* p, int64_t* num_incidences) {
  hm::clear_error();
  hm_status st = check_params(p);
  if (st != HM_OK) return st;
  if (!num_incidences) return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_static_size: null output");
  *num_incidences = draw_sizes(p, nullptr);
  return HM_OK;
}

hm_status hm_gen_static(const hm_gen_params* p, int64_t* offsets, int32_t* nodes,
                        int64_t nodes_capacity) {
  hm::clear_error();
  hm_status st = check_params(p);
  if (st != HM_OK) return st;
  if (!offsets || (!nodes && p->num_edges > 0))
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_static: null output buffer");
  std::vector<int32_t> sizes;
  int64_t total = draw_sizes(p, &sizes);
  if (nodes_capacity < total)
    return hm::fail(HM_ERR_BUFFER, "hm_gen_static: nodes_capacity below required incidences");

  // Zipf(s) popularity CDF over ranks 1..V.
  const int64_t V = p->num_nodes;
  std::vector<double> cdf((size_t)V);
  double acc = 0.0;
  for (int64_t r = 0; r < V; ++r) {
    acc += std::pow((double)(r + 1), -p->zipf_s);
    cdf[(size_t)r] = acc;
  }
  const double inv = 1.0 / acc;
  for (int64_t r = 0; r < V; ++r) cdf[(size_t)r] *= inv;
  cdf[(size_t)V - 1] = 1.0;

  hm::Xoshiro256ss rng(member_seed(p->seed));
  int64_t pos = 0;
  offsets[0] = 0;
  for (int64_t e = 0; e < p->num_edges; ++e) {
    const int32_t k = sizes[(size_t)e];
    int32_t* row = nodes + pos;
    for (int32_t j = 0; j < k; ++j) {
      for (;;) {
        double u = rng.uniform();
        int64_t r = (int64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (r >= V) r = V - 1;
        bool dup = false;
        for (int32_t t = 0; t < j; ++t)
          if (row[t] == (int32_t)r) { dup = true; break; }
        if (!dup) { row[j] = (int32_t)r; break; }
      }
    }
    pos += k;
    offsets[e + 1] = pos;
  }
  return HM_OK;
}

hm_status hm_gen_timestamps(int64_t n, double mean_gap_us, double burst_fraction,
                            int32_t burst_len, double burst_gap_us, uint64_t seed,
                            int64_t* times) {
  hm::clear_error();
  if (n < 0 || (n > 0 && !times)) return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_timestamps: bad n/out");
  if (!(mean_gap_us > 0.0) || !(burst_gap_us > 0.0) || burst_fraction < 0.0 || burst_fraction > 1.0)
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_timestamps: bad gap or fraction");
  if (burst_len < 1) return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_timestamps: burst_len < 1");
  // Units: n_burst bursts of burst_len edges plus singletons, in a seeded shuffled order.
  int64_t n_burst = (int64_t)std::llround((double)n * burst_fraction / (double)burst_len);
  if (n_burst * burst_len > n) n_burst = n / burst_len;
  const int64_t n_single = n - n_burst * burst_len;
  std::vector<uint8_t> unit((size_t)(n_single + n_burst), 0);
  for (int64_t i = 0; i < n_burst; ++i) unit[(size_t)i] = 1;
  hm::Xoshiro256ss rng(seed * 2ull + 0x3Dull);
  for (int64_t i = (int64_t)unit.size() - 1; i > 0; --i) {
    int64_t j = (int64_t)rng.below((uint64_t)(i + 1));
    std::swap(unit[(size_t)i], unit[(size_t)j]);
  }
  int64_t t = 0, k = 0;
  auto gap = [&](double mean) {
    double g = std::nearbyint(rng.exponential(mean));
    return g < 1.0 ? (int64_t)1 : (int64_t)g;
  };
  for (size_t u = 0; u < unit.size(); ++u) {
    t += gap(mean_gap_us);
    times[k++] = t;
    if (unit[u]) {
      for (int32_t b = 1; b < burst_len; ++b) {
        t += gap(burst_gap_us);
        times[k++] = t;
      }
    }
  }
  return HM_OK;
}

hm_status hm_gen_pick(int64_t n, int64_t k, uint64_t seed, int64_t* idx) {
  hm::clear_error();
  if (n < 0 || k < 0 || k > n || (k > 0 && !idx))
    return hm::fail(HM_ERR_BAD_INPUT, "hm_gen_pick: need 0 <= k <= n and an output buffer");
  // Partial Fisher-Yates over a sparse permutation: O(k) memory.
  std::vector<std::pair<int64_t, int64_t>> swaps;  // (position, value) overrides
  swaps.reserve((size_t)(2 * k));
  auto get = [&](int64_t pos) {
    for (auto it = swaps.rbegin(); it != swaps.rend(); ++it)
      if (it->first == pos) return it->second;
    return pos;
  };
  hm::Xoshiro256ss rng(seed * 2ull + 0x77ull);
  if (k <= 4096) {
    for (int64_t i = 0; i < k; ++i) {
      int64_t j = i + (int64_t)rng.below((uint64_t)(n - i));
      int64_t vi = get(i), vj = get(j);
      swaps.push_back({j, vi});
      swaps.push_back({i, vj});
      idx[i] = vj;
    }
    return HM_OK;
  }
  std::vector<int64_t> perm((size_t)n);
  for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = i;
  for (int64_t i = 0; i < k; ++i) {
    int64_t j = i + (int64_t)rng.below((uint64_t)(n - i));
    std::swap(perm[(size_t)i], perm[(size_t)j]);
    idx[i] = perm[(size_t)i];
  }
  return HM_OK;
}

}  // extern "C"
