// motif_update.hpp — replacement for dynamicHyperGraph/include/motif_update.hpp:1-21 over libhmotif.
// The delta is exact: counts(new) - counts(old) over the 30 MoCHy bins, computed on the GPU from
// the changed hyperedges only (the reference's anchor rule was not a difference of two counts).
#ifndef MOTIF_UPDATE_HPP
#define MOTIF_UPDATE_HPP

#include <vector>

#include "structure.hpp"

// deletedHyperedgeIds: 1-based ids live in the old snapshot; insertedHyperedgeIds: 1-based ids of
// the new snapshot (reused ids may appear in both lists); every other id must have the same row in
// both snapshots. outDeltaCounts receives 30 signed deltas (int, as in the reference).
void computeMotifCountsDelta(const CBSTContext& oldH2vCtx,
                             const CBSTContext& oldH2hCtx,
                             const CBSTContext& newH2vCtx,
                             const CBSTContext& newH2hCtx,
                             const std::vector<int>& deletedHyperedgeIds,
                             const std::vector<int>& insertedHyperedgeIds,
                             std::vector<int>& outDeltaCounts);

// Same, with 64-bit deltas.
void computeMotifCountsDelta(const CBSTContext& oldH2vCtx,
                             const CBSTContext& oldH2hCtx,
                             const CBSTContext& newH2vCtx,
                             const CBSTContext& newH2hCtx,
                             const std::vector<int>& deletedHyperedgeIds,
                             const std::vector<int>& insertedHyperedgeIds,
                             std::vector<long long>& outDeltaCounts);

#endif  // MOTIF_UPDATE_HPP
