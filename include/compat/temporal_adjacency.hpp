// temporal_adjacency.hpp — replacement for temporalDynamicHMotif/include/temporal_adjacency.hpp:1-55.
#ifndef TEMPORAL_ADJACENCY_HPP
#define TEMPORAL_ADJACENCY_HPP

#include <string>
#include <vector>

#include "temporal_structure.hpp"

struct TemporalHypergraphIndex {
    TemporalCBSTOperations h2v;
    TemporalCBSTOperations v2h;

    TemporalHypergraphIndex(int payloadCapacity, int alignment = 4)
        : h2v("H2V", payloadCapacity, alignment), v2h("V2H", payloadCapacity, alignment) {}

    void rotate() { h2v.rotate(); v2h.rotate(); }
    void resetNewest() { h2v.resetLayer(TemporalLayer::Newest); v2h.resetLayer(TemporalLayer::Newest); }
};

struct CSRNeighbors {
    int numRows = 0;
    std::vector<int> rowPtr;
    std::vector<int> colIdx;
};

struct CrossAdjacencySpec {
    TemporalLayer anchorLayer;
    TemporalLayer neighborLayer;
    std::vector<int> anchorHyperedgeIds;  // 1-based
};

// Neighbours across layers: row r lists the (1-based, ascending) hyperedges of neighborLayer whose
// H2V row shares a vertex with anchor anchorHyperedgeIds[r] of anchorLayer. (The reference left
// these as stubs returning empty adjacency.)
CSRNeighbors buildCrossAdjacency(const CrossAdjacencySpec& spec, const TemporalCBSTOperations& h2v,
                                 const TemporalCBSTOperations& v2h);
CSRNeighbors buildAdjOlderToMiddle(const std::vector<int>& anchorsMiddle, const TemporalHypergraphIndex& thg);
CSRNeighbors buildAdjMiddleToNewest(const std::vector<int>& anchorsMiddle, const TemporalHypergraphIndex& thg);
CSRNeighbors buildAdjOlderToNewest(const std::vector<int>& anchorsOlder, const TemporalHypergraphIndex& thg);

#endif  // TEMPORAL_ADJACENCY_HPP
