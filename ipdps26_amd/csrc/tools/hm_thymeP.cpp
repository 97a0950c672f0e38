// hm_thymeP — drop-in for THyMe+'s process contract (THyMe-main/code/main_thymeP.cpp:117-389):
//     hm_thymeP <dataset> <delta> [threads=8] [graph_file_override] [result_file_override]
// graph file defaults to ../data/<dataset>/<dataset>.txt and the result file to
// ../results/<dataset>_<delta as %f>_thymeP.txt; the result file holds the runtime on line 1
// and "i<TAB>count" for the 96 temporal h-motifs; stdout ends with TOTAL_RUNTIME_SECONDS.
// delta is read as a number in the timestamps' unit and must be integral here.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>

#include "hm_io.hpp"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::cerr << "Usage: " << argv[0]
              << " <dataset> <delta> [threads=8] [graph_file_override] [result_file_override]" << std::endl;
    return 1;
  }
  const std::string dataset = argv[1];
  const double delta_d = std::stod(argv[2]);
  std::string graph = "../data/" + dataset + "/" + dataset + ".txt";
  std::string result = "../results/" + dataset + "_" + std::to_string(delta_d) + "_thymeP.txt";
  if (argc >= 5) graph = argv[4];
  if (argc >= 6) result = argv[5];
  try {
    if (delta_d < 0 || delta_d != std::floor(delta_d)) throw std::runtime_error("delta must be a non-negative integer");
    hmotif::check(hm_init(0));
    std::vector<int64_t> times;
    hmotif::Csr csr = hmio::read_thyme_tsv(graph, &times);
    std::cout << "# of hyperedges:\t" << csr.size() << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    hmotif::TemporalGraph g(csr, times);
    hmotif::TemporalCounts c = hmotif::computeTemporalMotifCounts(g, (int64_t)delta_d);
    std::chrono::duration<double> run = std::chrono::steady_clock::now() - t0;
    std::printf("TOTAL_RUNTIME_SECONDS %f\n", run.count());
    FILE* out = std::fopen(result.c_str(), "w");
    if (!out) throw std::runtime_error("cannot write " + result);
    std::fprintf(out, "%f\n", run.count());
    for (int i = 0; i < HM_TEMPORAL_BINS; ++i) std::fprintf(out, "%d\t%llu\n", i + 1, (unsigned long long)c[i]);
    std::fclose(out);
  } catch (const std::exception& e) {
    std::cerr << "hm_thymeP: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
