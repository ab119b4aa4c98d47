// Group-by aggregation: device data structures and per-row operations.
//
// Table.  Open addressing over 2^k slots; a group's state row has the same index as its slot, so
// finding the slot is the whole assignment (no second indirection to a dense id).  Keys of at most
// 64 bits (one INT64/FLOAT64, one 32-bit lane with a null flag, two 32-bit lanes) live in one
// 64-bit cell claimed by compare-and-swap; a key equal to the empty marker and the null key of a
// single 64-bit lane use two reserved rows after the slots.  Wider keys are W 32-bit words: cell 0
// holds a tag (marker bit, 31 hash bits, word 0) claimed by CAS, cells 1..W-1 hold (1<<32 | word)
// written by the claiming lane; a reader that sees a cell still 0 retries the same slot on its next
// loop iteration, so no lane ever waits inside a branch for another lane of its wave.  Cells only
// ever go from empty to their final value, so a plain (possibly stale) load can only return the
// final value or "empty", and "empty" is confirmed by the CAS itself.
//
// State.  One row of `stride` 64-bit words per slot (array of structures: one group's update
// touches one or two cache lines), built so that a typical row costs few atomics and none of them
// needs its old value back:
//   ROWS                     rows of the group (COUNT(*)); word 0 of every row
//   NULLS(col)               null rows of `col`; COUNT(col) = ROWS - NULLS, touched only by nulls
//   SUM_I64(col)             sum of the low 32-bit halves (u64) and of the high halves (i64):
//                            the exact 128-bit total is hi * 2^32 + lo for groups below 2^32 rows,
//                            and a zero half (e.g. every non-negative value below 2^32) costs nothing
//   SUM_D128(col)            the same over the four 32-bit limbs of a DECIMAL128 lane
//   SUM_F64(col)             double sum
//   MIN/MAX_I32(col)         value as int64 with an impossible identity
//   MIN/MAX_I64(col)         order-preserving code; presence bit in the row's flags word
//   MIN/MAX_F64(col)         order-preserving code, first zero row (row<<1 | sign), first NaN row,
//                            first NaN's bits
//   FIRST_ROW                smallest global row index (first-seen emission order)
// A MIN/MAX row first compares against a plain read of the current best and skips the atomic
// when it cannot improve it (the best only moves one way, so a stale read is conservative).
//
// Update.  Workgroups walk tiles of rows.  An LDS table caches hot groups: keyed by the key itself
// for 64-bit keys (a hit needs no global access at all) and by slot for wide keys.  A row whose
// group has an LDS entry is folded in with LDS atomics; other rows update their global row
// directly.  Every few tiles the entries are flushed to the global rows (one set of atomics per
// entry), entries seen at least twice since the last flush keep their place, and the rest are
// evicted -- an adaptive hot-key cache for skewed (Zipf) keys.
#pragma once

#include "vr_device.h"
#include "vr_types.h"

namespace vr {

enum KeyMode : int32_t { KM_I64 = 0, KM_I32 = 1, KM_2X32 = 2, KM_WIDE = 3 };

enum OpCode : int32_t {
    OP_ROWS = 0,
    OP_NULLS = 1,
    OP_SUM_I64 = 2,
    OP_SUM_D128 = 3,
    OP_SUM_F64 = 4,
    OP_MIN_I32 = 5,
    OP_MAX_I32 = 6,
    OP_MIN_I64 = 7,
    OP_MAX_I64 = 8,
    OP_MIN_F64 = 9,
    OP_MAX_F64 = 10,
    OP_FIRST_ROW = 11,
    // 128-bit MIN/MAX (lo, hi, any): no 128-bit atomics on this part, so they are folded after the
    // update in row order over the rows sorted by slot (ordered_fold_kernel), never per row
    OP_MIN_D128 = 12,
    OP_MAX_D128 = 13
};

// counters[] words
enum { C_GROUPS = 0, C_RESERVED = 1, C_NAN = 2, C_ERROR = 3, C_OUT = 4, C_STOP = 5, C_POOL = 6, C_PSTOP = 7, C_WORDS = 8 };
enum { RES_SENTINEL = 1, RES_NULL = 2 };
enum {
    ERR_LONG_OVERFLOW = 1,
    ERR_DECIMAL_OVERFLOW = 2,
    ERR_AVG_OF_OVERFLOW = 4,
    ERR_TABLE_FULL = 8,
    ERR_SUM_OF_OVERFLOW = 16,  // a merged decimal sum that a partial left null or that overflowed
    ERR_WIDE_TOTAL = 32,       // an ungrouped final average whose exact total needs more than 128 bits
    ERR_INTERNAL = 64,         // the partitioned update's count pass and scatter disagreed (never expected)
    ERR_ORDER = 128            // the emission order's occupied slots differ from the group count (never expected)
};

// How two partial states of a group combine, word by word.
enum CombCode : uint8_t { CB_ADD = 0, CB_FADD = 1, CB_SMIN = 2, CB_SMAX = 3, CB_UMIN = 4, CB_UMAX = 5, CB_OR = 6 };

struct Op {
    int32_t code;
    int32_t col;   // value column, -1 for ROWS / FIRST_ROW
    int32_t off;   // first state word
    int16_t bit;   // presence bit (MIN/MAX_I64), else 0
    int16_t gate;  // BOOL value column: the op folds only rows where it is non-null and false
                   // (a merged decimal sum's isEmpty flag); -1: every row
};

struct KeySpec {
    int32_t mode;
    int32_t width;      // 32-bit words of a wide key (null mask word included)
    int32_t nkeys;
    int32_t null_word;  // index of the null-mask word, -1 when no key is nullable
    int32_t types[VR_MAX_KEYS];
    int32_t word_off[VR_MAX_KEYS];
};

// The slots are cut into 2^pbits regions of 2^rbits slots.  A key's region is the top pbits bits
// of its 64-bit hash and its home slot in the region the low rbits bits; probing wraps inside the
// region, so one workgroup can own a region (the partitioned update) while every other kernel sees
// an ordinary open-addressing table.
struct TableView {
    uint64_t* cells;
    uint64_t* state;
    uint64_t* counters;
    uint32_t* rgroups;  // groups per region (nullptr: not counted by this view's inserts)
    uint64_t mask;      // slots - 1
    uint64_t slots;     // reserved rows at [slots, slots + 2); row slots + 2 takes rows of a full region
    uint64_t rmask;     // region slots - 1
    int32_t pbits;
    int32_t rbits;
    int32_t width;      // cells per slot (1 for the 64-bit modes)
    int32_t stride;     // state words per row
};

__device__ __forceinline__ uint64_t region_of_hash(int pbits, uint64_t h) { return pbits ? (h >> (64 - pbits)) : 0; }
__device__ __forceinline__ uint64_t home_slot(const TableView& t, uint64_t h) {
    return (region_of_hash(t.pbits, h) << t.rbits) | (h & t.rmask);
}
__device__ __forceinline__ uint64_t next_slot(const TableView& t, uint64_t pos) {
    return (pos & ~t.rmask) | ((pos + 1) & t.rmask);
}

struct BatchDesc {
    DCol keys[VR_MAX_KEYS];
    DCol vals[VR_MAX_VALUES];
    const uint8_t* sel;
    int64_t n;
    int64_t row_base;    // global index of the batch's first row (first-seen order)
    int64_t tile_begin;  // first tile of the batch in a batched launch
};

struct UpdateArgs {
    TableView t;
    KeySpec ks;
    Op ops[VR_MAX_OPS];
    int32_t nops;
    int32_t flags_off;     // state word of the MIN/MAX_I64 presence bits, -1 when none
    int32_t lds_entries;   // power of two, 0: no LDS cache
    int32_t rows_per_thread;
    int32_t flush_tiles;   // tiles between LDS flushes
    int32_t filter_col;    // INT32 value column filtered by value < filter_threshold, -1: none
    int32_t filter_threshold;
    int32_t det;           // deterministic mode: double sums folded by the ordered pass
    uint32_t checked_sums; // deterministic mode: bit k = op k is a checked bigint sum, whose row-order
                           // prefixes the ordered pass checks (Math.addExact per row)
    int32_t num_batches;   // 1: `one` below; >1: `batches` (device array)
    int32_t epoch;         // update sequence number: the NaN flag holds the last epoch that saw one
    int32_t pad;
    int64_t tile_lo;       // tiles [tile_lo, total_tiles) of `batches` (or of `one`)
    int64_t total_tiles;
    // Load guard (resume != nullptr): before each tile a workgroup publishes its new groups and
    // stops once the table holds soft_limit groups, leaving in resume[blockIdx.x] the tile it would
    // have taken next (total_tiles or beyond when it ran out of tiles) and setting C_STOP.  A resumed
    // launch (resumed != 0, same grid) starts each workgroup at its resume entry.
    int64_t* resume;
    uint64_t soft_limit;
    int32_t resumed;
    int32_t pad2;
    uint64_t ident[64];    // identity value of every state word
    uint8_t comb[64];      // CombCode of every state word
    uint32_t* row_slot;    // deterministic: slot of every row of the batch (0xFFFFFFFF none)
    const BatchDesc* batches;
    BatchDesc one;
};

// Batch descriptor `i` of a device array that no kernel writes while it runs: read through the
// constant address space, so its fields come in scalar loads that the compiler may keep across
// the LDS atomics of a row loop.
__device__ __forceinline__ const BatchDesc& batch_at(const BatchDesc* batches, int64_t i) {
    const __attribute__((address_space(4))) BatchDesc* c = (const __attribute__((address_space(4))) BatchDesc*)batches;
    return *(const BatchDesc*)(c + i);
}

// The same for a descriptor every lane of the wave reads (a tile's batch): its address is made
// uniform, so the fields come in scalar loads once per tile rather than per row and operation.
__device__ __forceinline__ const BatchDesc& uniform_batch(const BatchDesc* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return batch_at((const BatchDesc*)(((uint64_t)hi << 32) | lo), 0);
}

__device__ __forceinline__ bool is_valid(const DCol& c, int64_t i) { return !c.valid || bit_at(c.valid, i); }

// A 32-bit key lane (INT32 / BOOL / UTF8 table-dictionary id).
__device__ __forceinline__ uint32_t key32(const DCol& c, int64_t i, bool* isnull) {
    *isnull = c.valid && !bit_at(c.valid, i);
    if (*isnull) return 0;
    switch (c.type) {
        case VR_BOOL: return bit_at((const uint8_t*)c.data, i) ? 1u : 0u;
        case VR_UTF8: {
            int32_t code = ld_stream((const int32_t*)c.data + i);
            if (c.remap) {
                int32_t id = (code >= 0 && code < c.remap_len) ? c.remap[code] : -1;
                if (id < 0) {  // a null dictionary entry: the row is null through the dictionary
                    *isnull = true;
                    return 0;
                }
                return (uint32_t)id;
            }
            return (uint32_t)code;
        }
        default: return (uint32_t)ld_stream((const int32_t*)c.data + i);
    }
}

__device__ __forceinline__ void mark_reserved(uint64_t* counters, uint64_t bit) {
    if ((counters[C_RESERVED] & bit) == 0) atomicOr((unsigned long long*)&counters[C_RESERVED], (unsigned long long)bit);
}

// A region with no empty slot left: the row goes to the spare row slots + 2 and the update fails
// with ERR_TABLE_FULL (the load limits keep this from happening to keys that hash evenly).
__device__ __forceinline__ uint64_t region_full(const TableView& t) {
    if ((t.counters[C_ERROR] & ERR_TABLE_FULL) == 0)
        atomicOr((unsigned long long*)&t.counters[C_ERROR], (unsigned long long)ERR_TABLE_FULL);
    return t.slots + 2;
}

__device__ __forceinline__ void count_insert(const TableView& t, uint64_t slot) {
    if (t.rgroups) atomicAdd(&t.rgroups[slot >> t.rbits], 1u);
}

// Slot of a 64-bit key, claiming an empty cell for a new key; `inserted` reports a claim.
__device__ __forceinline__ uint64_t find_or_insert64(const TableView& t, uint64_t key, bool* inserted) {
    *inserted = false;
    if (key == EMPTY_KEY) {
        mark_reserved(t.counters, RES_SENTINEL);
        return t.slots;
    }
    uint64_t pos = home_slot(t, fmix64(key));
    for (uint64_t probes = 0; probes <= t.rmask; probes++) {
        uint64_t c = t.cells[pos];
        if (c == key) return pos;
        if (c == EMPTY_KEY) {
            uint64_t old = atomic_cas_u64(&t.cells[pos], EMPTY_KEY, key);
            if (old == EMPTY_KEY) {
                *inserted = true;
                count_insert(t, pos);
                return pos;
            }
            if (old == key) return pos;
        }
        pos = next_slot(t, pos);
    }
    return region_full(t);
}

__device__ __forceinline__ uint64_t find64(const TableView& t, uint64_t key) {
    if (key == EMPTY_KEY) return t.slots;
    uint64_t pos = home_slot(t, fmix64(key));
    for (uint64_t probes = 0; probes <= t.rmask; probes++) {
        uint64_t c = t.cells[pos];
        if (c == key) return pos;
        if (c == EMPTY_KEY) {
            c = ld_coherent(&t.cells[pos]);
            if (c == key) return pos;
            if (c == EMPTY_KEY) return ~0ull;
        }
        pos = next_slot(t, pos);
    }
    return ~0ull;
}

__device__ __forceinline__ uint64_t hash_words(const uint32_t* w, int n) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)n;
    for (int i = 0; i < n; i++) h = fmix64(h ^ ((uint64_t)w[i] * 0xC2B2AE3D27D4EB4Full));
    return h;
}

// Slot of a wide key of `W` words.  Each loop iteration is one step for every lane still probing;
// a lane that found the tag but a later cell not yet written stays on the slot for the next step.
__device__ __forceinline__ uint64_t find_or_insert_wide(const TableView& t, const uint32_t* w, int W,
                                                        bool* inserted, bool insert) {
    *inserted = false;
    uint64_t h = hash_words(w, W);
    uint64_t tag = (1ull << 63) | (((h >> 32) & 0x7FFFFFFFull) << 32) | (uint64_t)w[0];
    uint64_t pos = home_slot(t, h);
    bool coherent = false;
    uint64_t probes = 0;
    while (true) {
        if (probes > t.rmask) return insert ? region_full(t) : ~0ull;
        uint64_t* s = t.cells + pos * (uint64_t)W;
        uint64_t c = coherent ? ld_coherent(s) : s[0];
        if (c == 0) {
            if (!insert) {
                c = ld_coherent(s);
                if (c == 0) return ~0ull;
            } else {
                uint64_t old = atomic_cas_u64(s, 0, tag);
                if (old == 0) {
                    for (int i = 1; i < W; i++) st_coherent(s + i, (1ull << 32) | (uint64_t)w[i]);
                    *inserted = true;
                    count_insert(t, pos);
                    return pos;
                }
                c = old;
            }
        }
        if (c == tag) {
            bool pending = false, same = true;
            for (int i = 1; i < W; i++) {
                uint64_t v = s[i];
                if (v == 0) v = ld_coherent(s + i);
                if (v == 0) {
                    pending = true;
                    break;
                }
                if (v != ((1ull << 32) | (uint64_t)w[i])) {
                    same = false;
                    break;
                }
            }
            if (pending) {  // the claiming lane's words are still in flight: look again next step
                coherent = true;
                continue;
            }
            if (same) return pos;
        }
        coherent = false;
        pos = next_slot(t, pos);
        probes++;
    }
}

// The 64-bit key of row i (64-bit key modes).  Returns false for a row that lives in a reserved
// row without a key (the null of a single 64-bit lane): then *slot is that row.
__device__ __forceinline__ bool key64_of(const UpdateArgs& a, const BatchDesc& b, int64_t i, uint64_t* key,
                                         uint64_t* slot) {
    switch (a.ks.mode) {
        case KM_I64: {
            const DCol& c = b.keys[0];
            if (c.valid && !bit_at(c.valid, i)) {
                *slot = a.t.slots + 1;
                return false;
            }
            *key = (uint64_t)ld_stream((const int64_t*)c.data + i);
            return true;
        }
        case KM_I32: {
            bool nul;
            uint32_t v = key32(b.keys[0], i, &nul);
            *key = nul ? (1ull << 32) : (uint64_t)v;
            return true;
        }
        default: {
            bool n0, n1;
            uint32_t v0 = key32(b.keys[0], i, &n0);
            uint32_t v1 = key32(b.keys[1], i, &n1);
            *key = ((uint64_t)v0 << 32) | (uint64_t)v1;
            return true;
        }
    }
}

__device__ __forceinline__ void wide_words(const UpdateArgs& a, const BatchDesc& b, int64_t i, uint32_t* w) {
    const KeySpec& ks = a.ks;
    for (int k = 0; k < ks.width; k++) w[k] = 0;
    uint32_t nm = 0;
    for (int c = 0; c < ks.nkeys; c++) {
        const DCol& col = b.keys[c];
        int o = ks.word_off[c];
        switch (ks.types[c]) {
            case VR_INT64:
            case VR_FLOAT64:
                if (col.valid && !bit_at(col.valid, i)) {
                    nm |= 1u << c;
                } else {
                    uint64_t v = (uint64_t)ld_stream((const int64_t*)col.data + i);
                    w[o] = (uint32_t)v;
                    w[o + 1] = (uint32_t)(v >> 32);
                }
                break;
            case VR_DECIMAL128:
                if (col.valid && !bit_at(col.valid, i)) {
                    nm |= 1u << c;
                } else {
                    const uint64_t* p = (const uint64_t*)col.data + 2 * i;
                    uint64_t lo = ld_stream(p), hi = ld_stream(p + 1);
                    w[o] = (uint32_t)lo;
                    w[o + 1] = (uint32_t)(lo >> 32);
                    w[o + 2] = (uint32_t)hi;
                    w[o + 3] = (uint32_t)(hi >> 32);
                }
                break;
            default: {
                bool nul;
                uint32_t v = key32(col, i, &nul);
                if (nul) nm |= 1u << c;
                else w[o] = v;
            }
        }
    }
    if (ks.null_word >= 0) w[ks.null_word] = nm;
}

// The state row of row i (global table), claiming a new group when `insert`; ~0 when absent.
__device__ __forceinline__ uint64_t row_slot_of(const UpdateArgs& a, const BatchDesc& b, int64_t i, bool* inserted,
                                                bool insert) {
    *inserted = false;
    if (a.ks.mode != KM_WIDE) {
        uint64_t key, slot;
        if (!key64_of(a, b, i, &key, &slot)) {
            if (insert) mark_reserved(a.t.counters, RES_NULL);
            return slot;
        }
        return insert ? find_or_insert64(a.t, key, inserted) : find64(a.t, key);
    }
    uint32_t w[4 * VR_MAX_KEYS + 1];
    wide_words(a, b, i, w);
    return find_or_insert_wide(a.t, w, a.ks.width, inserted, insert);
}

// ------------------------------------------------------------------ per-row updates

__device__ __forceinline__ int64_t read_long(const DCol& c, int64_t i) {
    return c.type == VR_INT32 ? (int64_t)ld_stream((const int32_t*)c.data + i) : ld_stream((const int64_t*)c.data + i);
}

__device__ __forceinline__ void add_if(uint64_t* w, uint64_t v) {
    if (v) atomic_add_u64(w, v);
}

// A row of an input batch, read from its columns.
struct ColSrc {
    const BatchDesc& b;
    int64_t i;
    __device__ __forceinline__ bool valid(int c) const { return is_valid(b.vals[c], i); }
    __device__ __forceinline__ int64_t lng(int c) const { return read_long(b.vals[c], i); }
    __device__ __forceinline__ int32_t i32(int c) const { return ld_stream((const int32_t*)b.vals[c].data + i); }
    __device__ __forceinline__ uint64_t u64(int c) const { return (uint64_t)ld_stream((const int64_t*)b.vals[c].data + i); }
    __device__ __forceinline__ void d128(int c, uint64_t* lo, uint64_t* hi) const {
        const uint64_t* p = (const uint64_t*)b.vals[c].data + 2 * i;
        *lo = ld_stream(p);
        *hi = ld_stream(p + 1);
    }
    __device__ __forceinline__ uint64_t row() const { return (uint64_t)(b.row_base + i); }
    __device__ __forceinline__ bool gate_open(int c) const {
        return is_valid(b.vals[c], i) && !bit_at((const uint8_t*)b.vals[c].data, i);
    }
};

// A value already in registers: validity, the lane's 64-bit word(s) (a DECIMAL128 in x0/x1, an
// INT32 in the low half of x0) and the global row index.
struct RegSrc {
    bool v;
    uint64_t x0, x1, gi;
    int32_t type;
    __device__ __forceinline__ bool valid(int) const { return v; }
    __device__ __forceinline__ uint64_t u64(int) const { return x0; }
    __device__ __forceinline__ int32_t i32(int) const { return (int32_t)(uint32_t)x0; }
    __device__ __forceinline__ int64_t lng(int) const { return type == VR_INT32 ? (int64_t)(int32_t)(uint32_t)x0 : (int64_t)x0; }
    __device__ __forceinline__ void d128(int, uint64_t* lo, uint64_t* hi) const {
        *lo = x0;
        *hi = x1;
    }
    __device__ __forceinline__ uint64_t row() const { return gi; }
    // Gated operations never reach a register source: specs that have one take the per-tile path.
    __device__ __forceinline__ bool gate_open(int) const { return true; }
};

// Folds one operation of one row into the state words at `row` (an LDS entry or a global row)
// with atomics.
template <class Src>
__device__ __forceinline__ void apply_one(const UpdateArgs& a, const Op& op, const Src& src, uint64_t* row) {
    if (op.gate >= 0 && !src.gate_open(op.gate)) return;
    {
        uint64_t* w = row + op.off;
        switch (op.code) {
            case OP_ROWS: atomic_add_u64(w, 1); break;
            case OP_NULLS:
                if (!src.valid(op.col)) atomic_add_u64(w, 1);
                break;
            case OP_SUM_I64: {
                if (!src.valid(op.col)) break;
                int64_t x = src.lng(op.col);
                add_if(w, (uint64_t)(uint32_t)x);
                add_if(w + 1, (uint64_t)(x >> 32));  // the signed high half
                break;
            }
            case OP_SUM_D128: {
                if (!src.valid(op.col)) break;
                uint64_t lo, hi;
                src.d128(op.col, &lo, &hi);
                add_if(w, lo & 0xFFFFFFFFull);
                add_if(w + 1, lo >> 32);
                add_if(w + 2, hi & 0xFFFFFFFFull);
                add_if(w + 3, (uint64_t)((int64_t)hi >> 32));
                break;
            }
            case OP_SUM_F64: {
                if (a.det || !src.valid(op.col)) break;
                atomic_add_f64(w, __longlong_as_double((long long)src.u64(op.col)));
                break;
            }
            case OP_MIN_I32:
            case OP_MAX_I32: {
                if (!src.valid(op.col)) break;
                int64_t x = src.i32(op.col);
                if (op.code == OP_MIN_I32) {
                    if (x < (int64_t)*w) atomic_min_i64(w, x);
                } else if (x > (int64_t)*w) {
                    atomic_max_i64(w, x);
                }
                break;
            }
            case OP_MIN_I64:
            case OP_MAX_I64: {
                if (!src.valid(op.col)) break;
                uint64_t o = src.u64(op.col) ^ 0x8000000000000000ull;
                uint64_t* f = row + a.flags_off;
                uint64_t bit = 1ull << op.bit;
                bool has = (*f & bit) != 0;
                if (op.code == OP_MIN_I64) {
                    if (!has || o < *w) atomic_min_u64(w, o);
                } else if (!has || o > *w) {
                    atomic_max_u64(w, o);
                }
                if (!has) atomicOr((unsigned long long*)f, bit);
                break;
            }
            case OP_MIN_F64:
            case OP_MAX_F64: {
                if (!src.valid(op.col)) break;
                uint64_t bits = src.u64(op.col);
                uint64_t o = order_f64(bits);
                if (op.code == OP_MIN_F64) {
                    if (o < *w) atomic_min_u64(w, o);
                } else if (o > *w) {
                    atomic_max_u64(w, o);
                }
                // Ties between +0.0 / -0.0 and between NaN payloads keep the first-seen value
                // (DoubleMinMax.offer replaces only on a strict improvement): record the first
                // zero row with its sign and the first NaN row.
                uint64_t gi = src.row();
                if (is_zero_bits(bits)) {
                    atomic_min_u64(w + 1, (gi << 1) | (bits >> 63));
                } else if (is_nan_bits(bits)) {
                    atomic_min_u64(w + 2, gi);
                    uint64_t* flag = &a.t.counters[C_NAN];
                    if (*flag != (uint64_t)a.epoch) atomic_max_u64(flag, (uint64_t)a.epoch);
                }
                break;
            }
            case OP_FIRST_ROW: {
                uint64_t gi = src.row();
                if (*w > gi) atomic_min_u64(w, gi);
                break;
            }
            case OP_MIN_D128:
            case OP_MAX_D128: break;  // folded in row order after the update
        }
    }
}

// Folds one row (every operation) into the state words at `row` with atomics.
template <class Src>
__device__ __forceinline__ void apply_src(const UpdateArgs& a, const Src& src, uint64_t* row) {
    for (int k = 0; k < a.nops; k++) apply_one(a, a.ops[k], src, row);
}

// Folds row i of batch b into the state words at `row` (an LDS entry or a global row).
__device__ __forceinline__ void apply_ops(const UpdateArgs& a, const BatchDesc& b, int64_t i, uint64_t* row) {
    apply_src(a, ColSrc{b, i}, row);
}

__device__ __forceinline__ uint64_t combine_word(uint8_t code, uint64_t acc, uint64_t x) {
    switch (code) {
        case CB_ADD: return acc + x;
        case CB_FADD: return (uint64_t)__double_as_longlong(__longlong_as_double((long long)acc) + __longlong_as_double((long long)x));
        case CB_SMIN: return (int64_t)x < (int64_t)acc ? x : acc;
        case CB_SMAX: return (int64_t)x > (int64_t)acc ? x : acc;
        case CB_UMIN: return x < acc ? x : acc;
        case CB_UMAX: return x > acc ? x : acc;
        default: return x | acc;
    }
}

// Combines a partial state into a row that no other thread updates meanwhile (plain stores): the
// row is read eight words at a time, so its loads are in flight together.
__device__ __forceinline__ void merge_plain(const UpdateArgs& a, const uint64_t* l, uint64_t* g, int S) {
    for (int k0 = 0; k0 < S; k0 += 8) {
        uint64_t y[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (k0 + j < S) y[j] = g[k0 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + j;
            if (k >= S) break;
            const uint64_t x = l[k];
            if (x != a.ident[k]) g[k] = combine_word(a.comb[k], y[j], x);
        }
    }
}

// Adds an LDS entry's partial state into the global row (one atomic per non-identity word).
__device__ __forceinline__ void flush_entry(const UpdateArgs& a, const uint64_t* e, uint64_t* g) {
    for (int k = 0; k < a.nops; k++) {
        const Op& op = a.ops[k];
        const uint64_t* l = e + op.off;
        uint64_t* w = g + op.off;
        switch (op.code) {
            case OP_ROWS:
            case OP_NULLS: add_if(w, l[0]); break;
            case OP_SUM_I64:
                add_if(w, l[0]);
                add_if(w + 1, l[1]);
                break;
            case OP_SUM_D128:
                for (int q = 0; q < 4; q++) add_if(w + q, l[q]);
                break;
            case OP_SUM_F64:
                if (!a.det && l[0]) atomic_add_f64(w, __longlong_as_double((long long)l[0]));
                break;
            case OP_MIN_I32:
                if ((int64_t)l[0] < (int64_t)*w) atomic_min_i64(w, (int64_t)l[0]);
                break;
            case OP_MAX_I32:
                if ((int64_t)l[0] > (int64_t)*w) atomic_max_i64(w, (int64_t)l[0]);
                break;
            case OP_MIN_I64:
            case OP_MAX_I64: {
                uint64_t f = e[a.flags_off] & (1ull << op.bit);
                if (f) {
                    if (op.code == OP_MIN_I64) atomic_min_u64(w, l[0]);
                    else atomic_max_u64(w, l[0]);
                    if ((g[a.flags_off] & f) == 0) atomicOr((unsigned long long*)&g[a.flags_off], f);
                }
                break;
            }
            case OP_MIN_F64:
            case OP_MAX_F64:
                if (op.code == OP_MIN_F64 ? l[0] < *w : l[0] > *w) {
                    if (op.code == OP_MIN_F64) atomic_min_u64(w, l[0]);
                    else atomic_max_u64(w, l[0]);
                }
                if (l[1] < w[1]) atomic_min_u64(w + 1, l[1]);
                if (l[2] < w[2]) atomic_min_u64(w + 2, l[2]);
                break;
            case OP_FIRST_ROW:
                if (l[0] < *w) atomic_min_u64(w, l[0]);
                break;
            case OP_MIN_D128:
            case OP_MAX_D128: break;  // folded in row order after the update
        }
    }
}

}  // namespace vr
