// motif.hpp — replacement for dynamicHyperGraph/include/motif.hpp:1-15 over libhmotif.
// computeMotifCounts counts every connected triple exactly as MoCHy-E does (the reference kernel
// counted closed triples only, into int32), on the GPU, and prints the same one-line summary.
#ifndef MOTIF_HPP
#define MOTIF_HPP

#include <vector>

#include "structure.hpp"

// Counts the hypergraph whose rows are the H2V rows 1..numHyperedges of h2vCtx (v2hCtx and h2hCtx
// are accepted for signature compatibility; adjacency is derived from H2V) and prints
// "Motif counts (30 bins): c0 ... c29".
void computeMotifCounts(const CBSTContext& h2vCtx,
                        const CBSTContext& v2hCtx,
                        const CBSTContext& h2hCtx,
                        int numHyperedges);

// Same count, returned (uint64 per bin) instead of printed.
void computeMotifCounts(const CBSTContext& h2vCtx,
                        const CBSTContext& v2hCtx,
                        const CBSTContext& h2hCtx,
                        int numHyperedges,
                        std::vector<unsigned long long>& outCounts30);

#endif  // MOTIF_HPP
