// structure.hpp — source-compatible replacement for the reference's CBST store interface
// (dynamicHyperGraph/include/structure.hpp:1-66), implemented over libhmotif.
//
// A driver written against the reference compiles unchanged with -I<repo>/include/compat and links
// -lhmotif_ref -lhmotif. The store keeps each key's row (hyperedge -> vertices, or vertex ->
// hyperedges) host-side; the counting entry points (motif.hpp, motif_update.hpp, temporal_count.hpp)
// build the device-resident graph from the H2V rows and count on the GPU. What differs from the
// reference, on purpose:
//   * rows are sets (a value added twice is kept once), read up to the first 0 or INT_MIN of the
//     padded flat payload exactly as the reference's kernels read them (kernel/device_utils.cuh:9-12);
//   * the device pointers of CBSTContext are null (there is no binary search tree); numRecords is
//     the largest key present, so ids 1..numRecords address the rows;
//   * errors throw hmotif_compat_error instead of calling exit(-1) (structure/operations.cu:24-29).
#ifndef STRUCTURE_HPP
#define STRUCTURE_HPP

#include <stdexcept>
#include <string>
#include <vector>

// Kept for source compatibility; the shim builds no tree.
struct CBSTNode {
    int index;
    int value;
    int length;
    int size;
    CBSTNode* left;
    CBSTNode* right;
    CBSTNode* parent;
};

struct CBSTContext {
    CBSTNode* d_nodes = nullptr;
    int* d_keys = nullptr;
    int* d_startOffsets = nullptr;
    int* d_flatPayload = nullptr;
    int* d_insertKeys = nullptr;
    int* d_insertPayload = nullptr;
    int* d_insertPrefixSizes = nullptr;
    int* d_relocationPlan = nullptr;
    int* d_avail = nullptr;
    int* d_subtreeAvail = nullptr;
    int fixedSize = 0;
    int numRecords = 0;
    int initialPayloadSize = 0;
    const char* datasetName = nullptr;
    int alignment = 4;
    void* store = nullptr;  // the shim's row store (owned by CBSTOperations)
};

class hmotif_compat_error : public std::runtime_error {
  public:
    explicit hmotif_compat_error(const std::string& m) : std::runtime_error(m) {}
};

// Host API (free functions), same argument meaning as the reference's: prefix sizes are inclusive
// running totals of the per-key payload lengths.
void constructCBST(int* keys, int* startOffsets, int numRecords, int* flatPayload, int flatPayloadSize, int payloadCapacity, const char* datasetName, CBSTContext& ctx);
void fillCBST(const std::vector<int>& insertKeys, const std::vector<int>& insertPayload, const std::vector<int>& insertPrefixSizes, CBSTContext& ctx);
void deleteCBST(const std::vector<int>& deleteKeys, CBSTContext& ctx);
void insertCBST(const std::vector<int>& newKeys, const std::vector<int>& newPayload, const std::vector<int>& newPrefixSizes, CBSTContext& ctx);
void unfillCBST(const std::vector<int>& keysToUnfill, const std::vector<int>& valuesToRemove, const std::vector<int>& removePrefixSizes, CBSTContext& ctx);

struct CBSTOperations {
    explicit CBSTOperations(const char* datasetName, int payloadCapacity, int alignment = 4);
    ~CBSTOperations();
    CBSTOperations(const CBSTOperations&) = delete;
    CBSTOperations& operator=(const CBSTOperations&) = delete;
    CBSTOperations(CBSTOperations&& other) noexcept;
    CBSTOperations& operator=(CBSTOperations&& other) noexcept;

    void construct(int* keys, int* startOffsets, int numRecords, int* flatPayload, int flatPayloadSize);
    void insert(const std::vector<int>& insertKeys, const std::vector<int>& insertPayload, const std::vector<int>& insertPrefixSizes);
    void erase(const std::vector<int>& deleteKeys);
    void findAndPrint(const std::vector<int>& ids) const;

    const CBSTContext& context() const;

  private:
    CBSTContext ctx_{};
    std::string name_;
    bool constructed_ = false;
};

// Row of `key` in a context (empty if absent); for drivers and tests that inspect the store.
std::vector<int> cbstRow(const CBSTContext& ctx, int key);

#endif  // STRUCTURE_HPP
