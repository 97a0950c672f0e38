// Motif numbering tables: 7-bit Venn-emptiness id -> 0-based bin (-1 = not a connected triple).
//
// The numbering is an interface the engine must keep (callers compare bins by position), so these
// tables carry exactly the values of the reference tables, shifted to 0-based bins:
//   kStaticBin   <- MoCHy id_to_index - 1   (MoCHy-master/motif_id.cpp:5-22;
//                   identical to dynamicHyperGraph/kernel/motif_utils.cuh:6-23)
//   kTemporalBin <- THyMe id_to_index_temporal - 1 (THyMe-main/code/motif_id.cpp:25-42)
// Venn id bits: 0..2 = first/second/third-only region, 3 = first∩second only,
// 4 = second∩third only, 5 = third∩first only, 6 = common to all three. For temporal bins the
// three hyperedges are taken in time order (earliest first), as THyMe's get_motif_index does
// (THyMe-main/code/motif_id.cpp:57-108).
#pragma once
#include <cstdint>

namespace hm {

#define HM_STATIC_BIN_TABLE { \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 0-7 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 8-15 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 16-23 */ \
    20, 22, 21, 23, 22, 24, 23, 25,  /* ids 24-31 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 32-39 */ \
    20, 21, 22, 23, 22, 23, 24, 25,  /* ids 40-47 */ \
    20, 22, 22, 24, 21, 23, 23, 25,  /* ids 48-55 */ \
    26, 27, 27, 28, 27, 28, 28, 29,  /* ids 56-63 */ \
    0,  1,  1,  2,  1,  2,  2,  3,   /* ids 64-71 */ \
    4,  5,  5,  7,  6,  8,  8,  9,   /* ids 72-79 */ \
    4,  6,  5,  8,  5,  8,  7,  9,   /* ids 80-87 */ \
    10, 12, 11, 13, 12, 14, 13, 15,  /* ids 88-95 */ \
    4,  5,  6,  8,  5,  7,  8,  9,   /* ids 96-103 */ \
    10, 11, 12, 13, 12, 13, 14, 15,  /* ids 104-111 */ \
    10, 12, 12, 14, 11, 13, 13, 15,  /* ids 112-119 */ \
    16, 17, 17, 18, 17, 18, 18, 19,  /* ids 120-127 */ \
  }

#define HM_TEMPORAL_BIN_TABLE { \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 0-7 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 8-15 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 16-23 */ \
    54, 60, 57, 66, 61, 72, 67, 75,  /* ids 24-31 */ \
    -1, -1, -1, -1, -1, -1, -1, -1,  /* ids 32-39 */ \
    55, 58, 62, 68, 63, 69, 73, 76,  /* ids 40-47 */ \
    56, 64, 65, 74, 59, 70, 71, 77,  /* ids 48-55 */ \
    78, 79, 80, 82, 81, 83, 84, 85,  /* ids 56-63 */ \
    95, 86, 87, 0,  88, 1,  2,  3,   /* ids 64-71 */ \
    89, 4,  5,  10, 92, 13, 14, 19,  /* ids 72-79 */ \
    90, 93, 6,  15, 7,  16, 11, 20,  /* ids 80-87 */ \
    22, 28, 25, 34, 29, 40, 35, 43,  /* ids 88-95 */ \
    91, 8,  94, 17, 9,  12, 18, 21,  /* ids 96-103 */ \
    23, 26, 30, 36, 31, 37, 41, 44,  /* ids 104-111 */ \
    24, 32, 33, 42, 27, 38, 39, 45,  /* ids 112-119 */ \
    46, 47, 48, 50, 49, 51, 52, 53,  /* ids 120-127 */ \
  }

constexpr int8_t kStaticBin[128] = HM_STATIC_BIN_TABLE;
constexpr int8_t kTemporalBin[128] = HM_TEMPORAL_BIN_TABLE;

// MoCHy prints 26 of the 30 static bins as "motif 1..26" (MoCHy-master/main_exact.cpp:133-137).
constexpr int kPrintedStaticBins[26] = {2,  3,  5,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16,
                                        17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29};

}  // namespace hm
