// temporal_count.hpp — replacement for temporalDynamicHMotif/include/temporal_count.hpp:1-27 over
// libhmotif (hm_count_strict_inc). These keep the reference's three-snapshot strict-increasing
// semantics, which are not THyMe's temporal h-motifs (for those: hm_count_temporal / hmotif.hpp).
#ifndef TEMPORAL_COUNT_HPP
#define TEMPORAL_COUNT_HPP

#include <string>
#include <vector>

#include "temporal_adjacency.hpp"

// Triples (i, j, k), i < j < k, of Older row i, Middle row j, Newest row k (rows 1..numRecords of
// each H2V layer), pairwise overlapping, into the 30 static bins — counted on the GPU with every
// candidate kept (the reference kernel dropped candidates beyond 1,024 per side).
void computeTemporalMotifCountsStrictInc(const TemporalHypergraphIndex& thg, std::vector<int>& outCounts30);
void computeTemporalMotifCountsStrictIncDelta(const TemporalHypergraphIndex& oldWindow,
                                              const TemporalHypergraphIndex& newWindow,
                                              std::vector<int>& outDeltaCounts30);

// Lines "A <hyperedgeId> v1 v2 ..." / "R <hyperedgeId> v1 v2 ..." add / remove vertices of the
// Newest H2V row and mirror the change into the Newest V2H rows. Returns false if the file cannot
// be read.
bool applyTemporalDeltasFromFile(const std::string& path, TemporalHypergraphIndex& thg);

#endif  // TEMPORAL_COUNT_HPP
