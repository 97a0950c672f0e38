// hm_exact_par — drop-in for MoCHy's exact counter process contract
// (MoCHy-master/main_exact_par.cpp:25-191):
//     hm_exact_par <num_threads> [graph_file]       (default graph_file: dblp_graph.txt)
// prints "# of nodes", "# of hyperedges", "motif n: c" for the 26 printed motifs and
// "TOTAL_RUNTIME_SECONDS x" (graph upload + projection + counting, as MoCHy times adjacency
// construction + counting). num_threads is accepted for compatibility; the count runs on GPU 0.
#include <chrono>
#include <cstdio>
#include <iostream>

#include "hm_io.hpp"

int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << " <num_threads> [graph_file]" << std::endl;
    return 1;
  }
  const std::string path = argc >= 3 ? argv[2] : "dblp_graph.txt";
  try {
    hmotif::check(hm_init(0));
    auto t0 = std::chrono::steady_clock::now();
    hmotif::Csr csr = hmio::read_mochy_csv(path);
    int32_t V = 0;
    for (int32_t v : csr.nodes) V = std::max(V, v + 1);
    std::cout << "# of nodes: " << V << '\n' << "# of hyperedges: " << csr.size() << '\n';
    std::chrono::duration<double> rd = std::chrono::steady_clock::now() - t0;
    std::cout << "Reading data done: " << rd.count() << " sec" << std::endl;
    auto t1 = std::chrono::steady_clock::now();
    hmotif::Graph g(csr);
    hmotif::StaticCounts c = hmotif::computeMotifCounts(g);
    std::chrono::duration<double> run = std::chrono::steady_clock::now() - t1;
    int index = 0;
    for (int i = 0; i < HM_STATIC_BINS; ++i) {
      if (i == 0 || i == 1 || i == 4 || i == 6) continue;
      std::cout << "motif " << ++index << ": " << c[i] << '\n';
    }
    std::cout << "\nTotal runtime: " << run.count() << std::endl;
    std::cout << "TOTAL_RUNTIME_SECONDS " << run.count() << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "hm_exact_par: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
