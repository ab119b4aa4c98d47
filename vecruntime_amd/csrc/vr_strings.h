// Device dictionary of byte strings (UTF8 key values -> dense int32 ids), shared by the group-by
// and the join.
//
// The reference maps each UTF8 key value to an id of the table's StringDictionary before it
// hashes the key (GroupKeyTable.toIds, kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java:
// 1308-1430; StringDictionary.indexOf): a group-by inserts unknown values, a join probe maps them
// to -1, which no build key carries.  Here the dictionary lives on the device, so plain UTF8
// columns in device memory are mapped without leaving it.
//
// Layout.  Open addressing over a power of two of 64-bit slots, at most half full: 0 is empty,
// else (fingerprint << 32) | ref, the fingerprint 31 bits of the string's 64-bit hash with the low
// bit set (never 0).  ref is the string's id -- or, only while one input is being inserted,
// PENDING | the input row that claimed the slot, so the row's own bytes (already in device memory)
// answer the comparisons of the other rows and no thread waits for another.  A second pass gives
// every claiming row an id (one counter), copies its bytes into the arena and rewrites its slot.
// Ids are dense from 0 in the order that pass meets the new strings, not first-seen order.
#pragma once

#include <vector>

#include "vr_internal.h"

namespace vr {

// The rows the string kernels read: row i is bytes[offsets[i], offsets[i + 1]), null where
// `valid` (nullable) says so.  Device pointers.
struct StrRows {
    const int32_t* offsets;
    const uint8_t* bytes;
    const uint8_t* valid;
    int64_t n;
    // [offsets[0], offsets[n]): set by StrTable::map, which keeps every byte read inside it
    int32_t begin = 0, end = 0;
};

class StrTable {
   public:
    // ids[i] (device, n int32) = the id of row i's string.  With `insert` unknown strings are
    // added; without, an unknown string gets -1.  A null row gets `null_id`.  The rows' bytes in
    // [offsets[0], offsets[n]) must total under 2^31.  Runs on `s` and returns once the ids are
    // written (it reads the table's counts back when inserting).
    vr_status map(Device* dev, const StrRows& rows, bool insert, int32_t null_id, int32_t* ids, hipStream_t s);
    // The same for a column in host memory (offsets and bytes copied to the device first).
    vr_status map_host(Device* dev, const int32_t* offsets, const uint8_t* bytes, const uint8_t* valid, int64_t n,
                       bool insert, int32_t null_id, int32_t* ids, hipStream_t s);
    int64_t size() const { return count_; }
    // The strings in id order, copied to the host: offsets[size + 1] and the bytes.
    vr_status read(std::vector<int32_t>* offsets, std::vector<uint8_t>* bytes, hipStream_t s) const;
    int64_t memory() const { return (int64_t)(cells_.bytes + spans_.bytes + arena_.bytes + scratch_.bytes); }

   private:
    vr_status reserve(Device* dev, int64_t more_strings, uint64_t more_bytes, hipStream_t s);
    DevBuf cells_, spans_, arena_, scratch_, counters_;
    uint64_t slots_ = 0;
    int64_t count_ = 0;         // ids given
    uint64_t arena_used_ = 0;   // bytes of the arena in use
};

}  // namespace vr
