// Sizes, record layout and column programs of the partitioned group-by update, shared by the
// kernels (vr_agg_part.h) and the host code that compiles them for a spec at run time.
#pragma once

#include "vr_types.h"

namespace vr {

constexpr int PART_BLK = 64;         // records per block: one per lane of a wave
constexpr int PART_REC_WORDS = 16;  // 32-bit words per record at most (a 64-byte record)
constexpr uint32_t NO_BLOCK = 0xFFFFFFFFu;
constexpr int PA_BLOCK = 1024;       // scatter workgroup
constexpr int PA_MAX_RPT = 2;        // rows per thread per scatter tile (generic kernel)
constexpr int PA_JIT_RPT = 4;        // the same, kernels compiled for their spec (fewer registers per row)
constexpr int PART_MAX_PROG = 4;     // value columns the partitioned update folds (programs)
constexpr int PB_BLOCK = 1024;       // merge workgroup
constexpr int PART_MAX_PBITS = 12;   // regions of the partitioned update: at most 2^12
constexpr int HS_BLOCK = 256;        // sampling workgroup
constexpr int HS_ROWS = 2048;        // rows per sample
constexpr int HS_SLOTS = 4096;       // LDS counting slots of a sample
constexpr int PB_CLAIM_BITS = 8192;  // merge: regions of at most this many slots claim new slots in LDS
constexpr int PB_FE = 2;             // merge cache entries per thread (the cache holds PB_FE * PB_BLOCK)
constexpr uint32_t HOT_EMPTY = 0xFFFFFFFFu;

struct RecLayout {
    int32_t words;                    // 32-bit words per record, a multiple of 4
    int32_t nvals;                    // value columns
    uint32_t vmask;                   // value columns whose validity the aggregates read
    int32_t row_off;                  // word offset of the global row index, -1: not carried
    int32_t valid_off;                // word offset of the validity bits (bit c: value column c)
    int32_t pad;
    int32_t val_off[VR_MAX_VALUES];   // word offset of each value lane carried, -1: not carried
    int32_t val_type[VR_MAX_VALUES];
};

// The aggregates of one value column, compiled by the host: which state words a row's value
// updates (offsets into the state row; -1 when absent).  Columns with a double MIN/MAX (first-seen
// zero and NaN rows) are folded op by op instead (CP_GENERIC).
enum ColProgFlags : uint16_t {
    CP_NULLS = 1,
    CP_SUM_I64 = 2,
    CP_SUM_D128 = 4,
    CP_SUM_F64 = 8,
    CP_MIN_I32 = 16,
    CP_MAX_I32 = 32,
    CP_MIN_I64 = 64,
    CP_MAX_I64 = 128,
    CP_GENERIC = 256
};

struct ColProg {
    uint16_t flags;
    int16_t col;
    int16_t type;
    int16_t nulls_off;
    int16_t sum_off;
    int16_t min_off;
    int16_t max_off;
    int16_t min_bit;  // presence bit of MIN/MAX_I64 in the flags word
    int16_t max_bit;
    int16_t rec_off;  // word offset of the lane in the record, -1: not carried
    int16_t width;    // bytes of the lane (4, 8 or 16); 0: validity only
    int16_t pad;
};

}  // namespace vr
