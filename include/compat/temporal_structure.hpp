// temporal_structure.hpp — replacement for temporalDynamicHMotif/include/temporal_structure.hpp:1-80
// over libhmotif: a three-slot ring of CBST stores (Older, Middle, Newest).
#ifndef TEMPORAL_STRUCTURE_HPP
#define TEMPORAL_STRUCTURE_HPP

#include <memory>
#include <string>
#include <vector>

#include "structure.hpp"

enum class TemporalLayer { Older = 0, Middle = 1, Newest = 2 };

class TemporalCBSTOperations {
  public:
    TemporalCBSTOperations(const std::string& label, int payloadCapacity, int alignment = 4);
    ~TemporalCBSTOperations();

    // Advances the ring: Middle becomes Older, Newest becomes Middle, and the slot that was Older
    // becomes Newest with its rows kept until the caller rebuilds it (as in the reference).
    void rotate();
    // As the reference: resets the layer's initialPayloadSize only; rows are not cleared.
    void resetLayer(TemporalLayer layer);

    void construct(TemporalLayer layer, int* cbstKeys, int* cbstStartOffsets, int numRecords,
                   int* flatPayload, int flatPayloadSize);
    void insert(TemporalLayer layer, const std::vector<int>& keys, const std::vector<int>& payload,
                const std::vector<int>& prefix);
    void erase(TemporalLayer layer, const std::vector<int>& keys);
    void fill(TemporalLayer layer, const std::vector<int>& keys, const std::vector<int>& payload,
              const std::vector<int>& prefix);
    void unfill(TemporalLayer layer, const std::vector<int>& keys,
                const std::vector<int>& valuesToRemove, const std::vector<int>& removePrefix);

    const CBSTContext& context(TemporalLayer layer) const;
    CBSTContext& context(TemporalLayer layer);

    // As the reference (temporalDynamicHMotif/src/temporal_structure.cpp:69-71): the physical slot
    // index of each logical layer, cast to TemporalLayer.
    TemporalLayer oldest() const;
    TemporalLayer middle() const;
    TemporalLayer newest() const;

  private:
    int resolveIndex(TemporalLayer layer) const;
    std::string label_;
    int oldestIdx_ = 0;
    std::unique_ptr<CBSTOperations> slots_[3];
};

#endif  // TEMPORAL_STRUCTURE_HPP
