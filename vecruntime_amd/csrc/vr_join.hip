// Hash join build and probe on gfx950: the vr_join_* entry points of include/vr.h.
//
// Replaces BuildTable.build (spark/src/main/scala/org/apache/spark/sql/vecruntime/VectorHashJoinExec.scala:
// 478-516), GroupKeyTable.lookup (kernels/.../GroupKeyTable.java:772-826) and the probe iterator's
// emitMatches / emitSemiAnti / gather (VectorHashJoinExec.scala:1038-1182,1605).
//
// Build table.  Buckets of 128 bytes: eight 64-bit keys, then eight 64-bit values, so a probe
// reads one cache line of keys and, on a hit, one value word of the same bucket.  A value is
// (rows of the key << 32) | payload, the payload being the build row itself for a unique key and
// the start of the key's run in `range_rows` (build rows ordered by key, build order within a key:
// the reference's head/next chains as a CSR) otherwise.  Keys of a single INT32/INT64 lane or two
// INT32 lanes; a key equal to the empty marker is kept in a side entry.
//
// Probe.  One pass over the streamed rows in tiles of 1024: each row's match count, a block scan,
// the tile's output offset by decoupled look-back over the preceding tiles (tile ids are taken from
// an atomic counter, so every earlier tile is already running), then the rows write their pairs
// or materialized columns.  The output order is therefore the reference's: streamed rows in order,
// each row's build rows in build order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>

#include "vr_device.h"
#include "vr_internal.h"
#include "vr_strings.h"

namespace vr {

constexpr int BUCKET = 8;  // cells per bucket
constexpr int PROBE_BLOCK = 256;
constexpr int PROBE_ITEMS = 8;  // rows per thread per tile
constexpr int64_t PROBE_TILE = (int64_t)PROBE_BLOCK * PROBE_ITEMS;
constexpr int MAX_PCOLS = 14;  // streamed column slots of a probe: 4 keys, 8 outputs, 2 residual operands
// In the value word of a bucket's first cell: an insert walked on past this (full) bucket, so a
// lookup that misses here must go on to the next bucket.  Row counts stay below 2^31, so the bit
// is free in every value word.
constexpr uint64_t PASSED = 1ull << 63;

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// A bucket is one 128-byte line of eight (key, value) cells, so a probe finds a key and its value
// with one line read.  The bucket count is not a power of two: the table is sized to its load
// limit (the home bucket is the hash's high half scaled to the bucket count), and probing goes on
// bucket by bucket, wrapping at the end.
struct JoinTable {
    uint64_t* buckets;  // nb * 16 words: key, value per cell
    uint64_t nb;        // buckets
    const uint32_t* range_rows;
    uint64_t sentinel_value;  // value of the key equal to EMPTY_KEY (0: absent)
    int32_t wide;             // cells hold (fingerprint << 32) | build row: keys compared in the build columns
    int32_t pad;
};

__host__ __device__ __forceinline__ uint64_t home_bucket(uint64_t key, uint64_t nb) { return ((fmix64(key) >> 32) * nb) >> 32; }
__host__ __device__ __forceinline__ uint64_t next_bucket(uint64_t b, uint64_t nb) { return b + 1 == nb ? 0 : b + 1; }

// ---- wide keys: more than 64 bits of key lanes (up to four columns, DECIMAL128 included)
//
// The table is keyed on a 64-bit hash of the key's lanes: a kernel hashes every build row into a
// 64-bit column, which the 64-bit build inserts as it would a key; a second kernel then checks
// that the rows of every cell hold one key (on a collision the build hashes again with another
// seed).  A probe looks its row's hash up like a 64-bit key and compares its lanes with the first
// build row of the run it found.  UTF8 key lanes hold the handle's dictionary ids, so equal
// strings are equal lanes.
constexpr int MAX_JOIN_KEYS = 4;

struct WideKeys {
    const uint8_t* valid[MAX_JOIN_KEYS];
    const void* data[MAX_JOIN_KEYS];
    int32_t width[MAX_JOIN_KEYS];  // bytes per row: 4, 8 or 16
    int32_t nk;
    int32_t pad;
    uint64_t seed;
};

__host__ __device__ __forceinline__ int key_lane_width(int32_t type) {
    return type == VR_INT64 || type == VR_FLOAT64 ? 8 : type == VR_DECIMAL128 ? 16 : 4;
}

__device__ __forceinline__ uint64_t wide_mix(uint64_t h, uint64_t w) { return fmix64(h ^ (w * 0xC2B2AE3D27D4EB4Full)); }

// The key's hash from its lanes at row i (data[c] + i * width[c]).
__device__ __forceinline__ uint64_t wide_hash(const void* const* data, const int32_t* width, int nk, uint64_t seed, int64_t i) {
    uint64_t h = seed + (uint64_t)nk;
    for (int c = 0; c < nk; c++) {
        if (width[c] == 4) {
            h = wide_mix(h, (uint64_t)ld_stream((const uint32_t*)data[c] + i));
        } else {
            const uint64_t* p = (const uint64_t*)data[c] + (width[c] == 16 ? 2 * i : i);
            h = wide_mix(h, ld_stream(p));
            if (width[c] == 16) h = wide_mix(h, ld_stream(p + 1));
        }
    }
    return h;
}

// Row i of lanes `a` equals row j of lanes `b` (same widths).
__device__ __forceinline__ bool wide_equal(const void* const* a, int64_t i, const void* const* b, int64_t j, const int32_t* width,
                                           int nk) {
    for (int c = 0; c < nk; c++) {
        if (width[c] == 4) {
            if (((const uint32_t*)a[c])[i] != ((const uint32_t*)b[c])[j]) return false;
        } else if (width[c] == 8) {
            if (((const uint64_t*)a[c])[i] != ((const uint64_t*)b[c])[j]) return false;
        } else {
            const uint64_t* x = (const uint64_t*)a[c] + 2 * i;
            const uint64_t* y = (const uint64_t*)b[c] + 2 * j;
            if (x[0] != y[0] || x[1] != y[1]) return false;
        }
    }
    return true;
}

// hashes[i] = the hash of build row i; valid: bit i set when no key lane of row i is null (each
// thread writes one 32-row word).
__global__ void wide_hash_kernel(WideKeys k, int64_t n, uint64_t* hashes, uint32_t* valid) {
    const int64_t words = (n + 31) / 32;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (int64_t)gridDim.x * blockDim.x) {
        uint32_t bits = 0;
        for (int64_t i = w * 32; i < n && i < w * 32 + 32; i++) {
            bool ok = true;
            for (int c = 0; c < k.nk; c++)
                if (k.valid[c] && !bit_at(k.valid[c], i)) ok = false;
            hashes[i] = ok ? wide_hash(k.data, k.width, k.nk, k.seed, i) : 0;
            if (ok) bits |= 1u << (i & 31);
        }
        valid[w] = bits;
    }
}

// Counts the build rows whose key differs from the first row of their cell's run (a hash
// collision): row_cell from the insert, runs in range_rows from cell_start; the rows of the key
// equal to EMPTY_KEY (cell 0xFFFFFFFE) run from `sentinel_start`.
__global__ void wide_verify_kernel(WideKeys k, int64_t n, const uint32_t* row_cell, const uint32_t* cell_start,
                                   const uint32_t* range_rows, uint32_t sentinel_start, uint64_t* collisions) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t cell = row_cell[i];
        if (cell == 0xFFFFFFFFu) continue;  // a null key
        const uint32_t first = range_rows[cell == 0xFFFFFFFEu ? sentinel_start : cell_start[cell]];
        if (!wide_equal(k.data, i, k.data, (int64_t)first, k.width, k.nk)) atomic_add_u64(collisions, 1);
    }
}

__device__ __forceinline__ bool key64_of(const DCol* k, int nk, int64_t i, uint64_t* key) {
    if (nk == 1) {
        const DCol& c = k[0];
        if (c.valid && !bit_at(c.valid, i)) return false;
        if (key_lane_width(c.type) == 8) *key = (uint64_t)ld_stream((const int64_t*)c.data + i);
        else *key = (uint64_t)(uint32_t)ld_stream((const int32_t*)c.data + i);
        return true;
    }
    if ((k[0].valid && !bit_at(k[0].valid, i)) || (k[1].valid && !bit_at(k[1].valid, i))) return false;
    uint32_t a = (uint32_t)ld_stream((const int32_t*)k[0].data + i);
    uint32_t b = (uint32_t)ld_stream((const int32_t*)k[1].data + i);
    *key = ((uint64_t)a << 32) | b;
    return true;
}

// Build: claims (or finds) the key's cell; returns the global cell index.  A key that goes on past
// a full bucket marks that bucket passed.
__global__ void join_insert_kernel(JoinTable t, DCol k0, DCol k1, int nk, int64_t n, uint32_t* row_cell,
                                   uint32_t* cell_count, uint64_t* sentinel_rows, uint32_t* passed) {
    DCol k[2] = {k0, k1};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key;
        if (!key64_of(k, nk, i, &key)) {  // a null key never matches (nonNullKeys)
            row_cell[i] = 0xFFFFFFFFu;
            continue;
        }
        if (key == EMPTY_KEY) {
            row_cell[i] = 0xFFFFFFFEu;
            atomic_add_u64(sentinel_rows, 1);
            continue;
        }
        uint64_t b = home_bucket(key, t.nb);
        uint32_t cell = 0xFFFFFFFFu;
        while (true) {
            uint64_t* cells = t.buckets + b * 16;
            for (int j = 0; j < BUCKET; j++) {
                uint64_t c = cells[2 * j];
                if (c == EMPTY_KEY) {
                    c = atomic_cas_u64(&cells[2 * j], EMPTY_KEY, key);
                    if (c == EMPTY_KEY) c = key;
                }
                if (c == key) {
                    cell = (uint32_t)(b * BUCKET + j);
                    break;
                }
            }
            if (cell != 0xFFFFFFFFu) break;
            atomicOr(&passed[b >> 5], 1u << (b & 31));  // every cell holds another key: go on
            b = next_bucket(b, t.nb);
        }
        row_cell[i] = cell;
        atomicAdd(&cell_count[cell], 1u);
    }
}

// Writes each occupied cell's value from its row count and run start, and the passed flags.
__global__ void join_values_kernel(JoinTable t, const uint32_t* cell_count, const uint32_t* cell_start, uint64_t ncells,
                                   const uint32_t* range_rows, const uint32_t* passed) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t cnt = cell_count[c];
        if (!cnt) continue;
        uint32_t start = cell_start[c];
        uint64_t v = ((uint64_t)cnt << 32) | (cnt == 1 ? range_rows[start] : start);
        const uint64_t b = c / BUCKET;
        if (c % BUCKET == 0 && ((passed[b >> 5] >> (b & 31)) & 1u)) v |= PASSED;  // a passed bucket is full
        t.buckets[b * 16 + 2 * (c % BUCKET) + 1] = v;
    }
}

// ---- cells that carry the build row's payload (unique keys, fixed-width payload lanes, no nulls)
//
// For inner and outer joins that write build columns, a second table whose 32-byte cells hold the
// key, the build row and up to 20 bytes of that row's payload lanes packed in column order, four
// cells to a 128-byte bucket: the line a probe reads to find the key also holds every build column
// it writes, so an output row costs no gather (a random line per build column otherwise).  Cell
// words: key; row (low half, bit 31 of a bucket's first cell = the passed flag) | payload word 0;
// payload words 1, 2; payload words 3, 4.  Home buckets and probing follow the main table.
constexpr int PBUCKET = 4;       // cells per payload bucket
constexpr int PAY_WORDS = 5;     // 32-bit payload words per cell
constexpr int PAY_MAX_COLS = 8;  // payload lanes of one table
constexpr uint32_t PROW_PASSED = 1u << 31;

struct PayTable {
    uint64_t* buckets;  // nb * 16 words
    uint64_t nb;
};

struct PaySrc {
    const void* data[PAY_MAX_COLS];
    int32_t width[PAY_MAX_COLS];  // 4, 8 or 16 bytes
    int32_t off[PAY_MAX_COLS];    // byte offset in the payload, a multiple of 4
    int32_t ncols;
    int32_t pad;
};

__global__ void pay_fill_kernel(uint64_t* buckets, uint64_t words) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        buckets[w] = (w & 3) == 0 ? EMPTY_KEY : 0ull;
}

// Inserts build row i (keys unique, none equal to the empty marker): the claiming lane writes the
// row (OR-ed in, as other rows' inserts may set the passed flag in the same word) and the payload.
__global__ void pay_insert_kernel(PayTable t, DCol k0, DCol k1, int nk, int64_t n, PaySrc src) {
    DCol k[2] = {k0, k1};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key;
        if (!key64_of(k, nk, i, &key)) continue;  // a null key never matches
        uint32_t pw[PAY_WORDS] = {0, 0, 0, 0, 0};
        for (int c = 0; c < src.ncols; c++) {
            const int wd = src.width[c] / 4;
            const uint32_t* from = (const uint32_t*)src.data[c] + (uint64_t)i * wd;
            for (int q = 0; q < wd; q++) pw[src.off[c] / 4 + q] = from[q];
        }
        uint64_t b = home_bucket(key, t.nb);
        while (true) {
            uint64_t* cells = t.buckets + b * 16;
            int claimed = -1;
            for (int j = 0; j < PBUCKET && claimed < 0; j++)
                if (cells[4 * j] == EMPTY_KEY && atomic_cas_u64(&cells[4 * j], EMPTY_KEY, key) == EMPTY_KEY) claimed = j;
            if (claimed >= 0) {
                uint64_t* c = cells + 4 * claimed;
                atomicOr((unsigned int*)&c[1], (unsigned int)i);
                ((uint32_t*)&c[1])[1] = pw[0];
                c[2] = (uint64_t)pw[1] | ((uint64_t)pw[2] << 32);
                c[3] = (uint64_t)pw[3] | ((uint64_t)pw[4] << 32);
                break;
            }
            atomicOr((unsigned int*)&cells[1], PROW_PASSED);  // every cell holds another key: go on
            b = next_bucket(b, t.nb);
        }
    }
}

// The payload cell of `key` walking from bucket b: its row (or -1) and payload words.
__device__ __forceinline__ int64_t pay_lookup_from(const PayTable& t, uint64_t key, uint64_t b, uint32_t (&pay)[PAY_WORDS]) {
    while (true) {
        const uint64_t* cells = t.buckets + b * 16;
        for (int j = 0; j < PBUCKET; j++)
            if (cells[4 * j] == key) {
                const uint64_t w1 = cells[4 * j + 1], w2 = cells[4 * j + 2], w3 = cells[4 * j + 3];
                pay[0] = (uint32_t)(w1 >> 32);
                pay[1] = (uint32_t)w2;
                pay[2] = (uint32_t)(w2 >> 32);
                pay[3] = (uint32_t)w3;
                pay[4] = (uint32_t)(w3 >> 32);
                return (int64_t)((uint32_t)w1 & ~PROW_PASSED);
            }
        if (!((uint32_t)cells[1] & PROW_PASSED)) return -1;
        b = next_bucket(b, t.nb);
    }
}

// Wave-cooperative lookup of the wave's 64 rows in the payload table: eight lanes read one bucket
// (lane j: half j & 1 of cell j >> 1), a ballot over the key halves finds the cell, shuffles hand
// its row and payload words to the row's lane.  A row whose home bucket lacks the key but was
// passed goes on by itself.  Returns the build row, -1 when absent (or `active` false).
__device__ __forceinline__ int64_t coop_pay_lookup(const PayTable& t, bool active, uint64_t key, uint32_t (&pay)[PAY_WORDS]) {
    const int lane = threadIdx.x & 63, g = lane >> 3, j = lane & 7;
    const uint32_t home = active ? (uint32_t)home_bucket(key, t.nb) : 0u;
    const u64x2* q = (const u64x2*)t.buckets;
    u64x2 x[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const uint32_t bb = __shfl(home, g * 8 + s);
        x[s] = q[(uint64_t)bb * 8 + j];
    }
    int64_t row = -1;
    bool cont = false;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const uint64_t kk = __shfl(key, g * 8 + s);
        const uint64_t mb = __ballot((j & 1) == 0 && x[s].x == kk);
        const uint64_t pb = __ballot(j == 0 && ((uint32_t)x[s].y & PROW_PASSED));
        const uint32_t gm = (uint32_t)(mb >> (g * 8)) & 0xFFu;
        const int m = gm ? __ffs(gm) - 1 : 0;
        const uint64_t a1 = __shfl(x[s].y, g * 8 + m);
        const uint64_t b0 = __shfl(x[s].x, g * 8 + m + 1);
        const uint64_t b1 = __shfl(x[s].y, g * 8 + m + 1);
        if (j == s) {
            row = gm ? (int64_t)((uint32_t)a1 & ~PROW_PASSED) : -1;
            cont = !gm && ((pb >> (g * 8)) & 1ull);
            pay[0] = (uint32_t)(a1 >> 32);
            pay[1] = (uint32_t)b0;
            pay[2] = (uint32_t)(b0 >> 32);
            pay[3] = (uint32_t)b1;
            pay[4] = (uint32_t)(b1 >> 32);
        }
    }
    if (active && cont) row = pay_lookup_from(t, key, next_bucket(home, t.nb), pay);
    return active ? row : -1;
}

// The key's value word (0 when absent), walking from bucket b.
__device__ __forceinline__ uint64_t lookup_from(const JoinTable& t, uint64_t key, uint64_t b) {
    const u64x2* cells = (const u64x2*)t.buckets;
    while (true) {
        u64x2 x[BUCKET];
#pragma unroll
        for (int j = 0; j < BUCKET; j++) x[j] = cells[b * BUCKET + j];
#pragma unroll
        for (int j = 0; j < BUCKET; j++)
            if (x[j].x == key) return x[j].y & ~PASSED;
        if (!(x[0].y & PASSED)) return 0;  // no key of this home or before went on past here
        b = next_bucket(b, t.nb);
    }
}

__device__ __forceinline__ uint64_t lookup(const JoinTable& t, uint64_t key) {
    if (key == EMPTY_KEY) return t.sentinel_value;
    return lookup_from(t, key, home_bucket(key, t.nb));
}

// Wave-cooperative lookup of the 64 rows the wave's lanes hold.  Eight lanes read one bucket, a
// 16-byte cell each, so a load instruction covers eight rows (a line each) and eight of them cover
// the wave; a ballot finds the matching cell and the bucket's passed flag, and a shuffle hands the
// matching cell's value to the row's lane.  A row whose home bucket lacks the key but was passed
// goes on bucket by bucket on its own lane (a few rows in a wave: looping the cooperative steps
// for them instead holds every round of the tile back and measured slower).  Lanes with `active`
// false read bucket 0 and get 0.  The whole wave must reach the call.
__device__ __forceinline__ uint64_t coop_lookup(const JoinTable& t, bool active, uint64_t key) {
    const int lane = threadIdx.x & 63, g = lane >> 3, j = lane & 7;
    const uint32_t home = active ? (uint32_t)home_bucket(key, t.nb) : 0u;
    const u64x2* cells = (const u64x2*)t.buckets;
    u64x2 x[BUCKET];
#pragma unroll
    for (int s = 0; s < BUCKET; s++) {
        const uint32_t bb = __shfl(home, g * 8 + s);
        x[s] = cells[(uint64_t)bb * BUCKET + j];
    }
    uint64_t res = 0;
    bool cont = false;
#pragma unroll
    for (int s = 0; s < BUCKET; s++) {
        const uint64_t kk = __shfl(key, g * 8 + s);
        const uint64_t mb = __ballot(x[s].x == kk);
        const uint64_t pb = __ballot(j == 0 && (x[s].y & PASSED));
        const uint32_t gm = (uint32_t)(mb >> (g * 8)) & 0xFFu;
        const uint64_t val = __shfl(x[s].y, g * 8 + (gm ? __ffs(gm) - 1 : 0));
        if (j == s) {
            res = gm ? (val & ~PASSED) : 0;
            cont = !gm && ((pb >> (g * 8)) & 1ull);
        }
    }
    if (active && cont) res = lookup_from(t, key, next_bucket(home, t.nb));
    return active ? res : 0;
}

struct OutCol {
    uint8_t* valid;  // nullptr: not written
    void* data;
    int32_t type;
    int32_t pad;
};

// The direct index over a single key's range [min, min + range): the key's value word, 0 when no
// build row has it (DenseKeyIndex.slots, kernels/src/main/java/io/vecruntime/kernels/DenseKeyIndex.java:44-110).
struct DenseIndex {
    const uint64_t* slots;  // nullptr: probe the hash table
    int64_t min;
    uint64_t range;
};

// The key's signed value (INT32 keys are read zero-extended into the 64-bit key).
__host__ __device__ __forceinline__ int64_t key_value(int32_t type, uint64_t key) {
    return type == VR_INT32 ? (int64_t)(int32_t)(uint32_t)key : (int64_t)key;
}

// DenseKeyIndex.lookup (:131-176): an offset outside the range, compared unsigned, is a miss.
__device__ __forceinline__ uint64_t dense_lookup(const DenseIndex& d, int32_t type, uint64_t key) {
    const uint64_t off = (uint64_t)key_value(type, key) - (uint64_t)d.min;
    return off < d.range ? d.slots[off] : 0;
}

// A streamed column of one batch.
struct PCol {
    const uint8_t* valid;  // nullptr: no nulls
    const void* data;
};

// One streamed batch of a probe launch.
struct PBatch {
    int64_t n;      // rows
    int64_t tile0;  // the launch's first tile of this batch
    const uint8_t* sel;
    int64_t reserved;
};

// A probe launch over one batch (the batch in the arguments) or many (the batch table in device
// memory).  A batch's streamed columns fill slots: the keys, then the columns to materialize, then
// the streamed operands of a residual.
struct ProbeArgs {
    JoinTable t;
    PayTable pt;                // buckets != nullptr: look keys up in the payload table (pay_off)
    int32_t pay_off[8];         // build output c: its byte offset in the payload; -1: not taken from it
    int32_t pay_words;          // payload words the table's cells use (staged in LDS per row)
    int32_t pad_pw;
    WideKeys bk;  // wide keys: the build side's key lanes
    DenseIndex dense;
    int32_t nk;
    int32_t join_type;
    int32_t drop_null_keys;  // null-aware anti: a null key counts as matched
    int32_t build_rows;      // 0: an empty build side
    int32_t ncb;             // slots per batch
    int32_t nprobe_out, nbuild_out;
    int32_t ctype[MAX_PCOLS];      // lane type of each slot
    int32_t build_from_key[8];     // build output c of a matched pair = key slot build_from_key[c] of the
                                   // streamed row (equal keys are the same bits); -1: gathered by build row
    DCol build_src[8];
    OutCol outs[16];
    int64_t capacity;
    int32_t* probe_idx;
    int32_t* build_idx;
    uint64_t* tile_state;    // per tile: (status << 62) | value, zeroed before the launch
    uint64_t* tile_counter;  // [0] next tile id, [1] total rows out
    int64_t ntiles;
    uint32_t* matched;       // build rows paired so far (one bit each), nullptr: not tracked
    // residual condition on the pairs (res_on 0: none): left + loff <op> right + roff, each operand
    // a streamed column (a slot, indexed by the streamed row) or a build column (by the build row)
    int32_t res_on, res_op;
    int32_t res_lb, res_rb;
    int32_t res_lslot, res_rslot;
    DCol res_lbuild, res_rbuild;
    int64_t res_loff, res_roff;
    const PBatch* batches;      // many batches: the batch table
    const PCol* bcols;          // many batches: batch b's slots at [b * ncb, (b + 1) * ncb)
    const int32_t* tile_batch;  // many batches: the batch of each tile
    int64_t* batch_end;         // many batches: output rows through each batch's last tile
    PBatch one;                 // one batch
    PCol one_cols[MAX_PCOLS];
};

template <bool MULTI>
__device__ __forceinline__ PCol slot_col(const ProbeArgs& a, const PCol* bc, int c) {
    return MULTI ? bc[c] : a.one_cols[c];
}

// The 64-bit probe key of row i from the key slots; false when a key is null.
__device__ __forceinline__ bool probe_key(const PCol& c0, const PCol& c1, int32_t t0, int nk, int64_t i, uint64_t* key) {
    if (nk == 1) {
        if (c0.valid && !bit_at(c0.valid, i)) return false;
        if (key_lane_width(t0) == 8) *key = (uint64_t)ld_stream((const int64_t*)c0.data + i);
        else *key = (uint64_t)(uint32_t)ld_stream((const int32_t*)c0.data + i);
        return true;
    }
    if ((c0.valid && !bit_at(c0.valid, i)) || (c1.valid && !bit_at(c1.valid, i))) return false;
    const uint32_t a = (uint32_t)ld_stream((const int32_t*)c0.data + i);
    const uint32_t b = (uint32_t)ld_stream((const int32_t*)c1.data + i);
    *key = ((uint64_t)a << 32) | b;
    return true;
}

// One operand of the residual plus its offset in the lane's width (wrapping); false when null.
__device__ __forceinline__ bool res_operand(const DCol& c, int64_t row, int64_t off, int64_t* v) {
    if (c.valid && !bit_at(c.valid, row)) return false;
    if (c.type == VR_INT32) *v = (int32_t)((uint32_t)((const int32_t*)c.data)[row] + (uint32_t)(int32_t)off);
    else *v = (int64_t)((uint64_t)((const int64_t*)c.data)[row] + (uint64_t)off);
    return true;
}

// PairPredicate.test (VectorHashJoinExec.scala:1287-1303) for streamed row i and build row b.
__device__ __forceinline__ bool res_passes(const ProbeArgs& a, const DCol& l, const DCol& r, int64_t i, int64_t b) {
    int64_t x, y;
    if (!res_operand(l, a.res_lb ? b : i, a.res_loff, &x) || !res_operand(r, a.res_rb ? b : i, a.res_roff, &y)) return false;
    const int c = (x > y) - (x < y);
    switch (a.res_op) {
        case VR_CMP_EQ: return c == 0;
        case VR_CMP_NE: return c != 0;
        case VR_CMP_LT: return c < 0;
        case VR_CMP_LE: return c <= 0;
        case VR_CMP_GT: return c > 0;
        default: return c >= 0;
    }
}

// Candidate k of a key's value word: the single row, or the k-th row of its run in build order.
__device__ __forceinline__ int64_t candidate(const ProbeArgs& a, uint32_t matches, uint32_t payload, uint32_t k) {
    return matches == 1 ? (int64_t)payload : (int64_t)a.t.range_rows[payload + k];
}

__device__ __forceinline__ void copy_value(const DCol& src, int64_t i, const OutCol& dst, int64_t o) {
    if (i < 0) {  // the padded side of an outer pair: null (validity words start cleared), data zero
        switch (dst.type) {
            case VR_INT32: ((int32_t*)dst.data)[o] = 0; break;
            case VR_INT64:
            case VR_FLOAT64: ((int64_t*)dst.data)[o] = 0; break;
            case VR_DECIMAL128:
                ((int64_t*)dst.data)[2 * o] = 0;
                ((int64_t*)dst.data)[2 * o + 1] = 0;
                break;
            default: break;
        }
        return;
    }
    if (dst.valid && (!src.valid || bit_at(src.valid, i))) atomicOr((unsigned int*)dst.valid + (o >> 5), 1u << (o & 31));
    switch (src.type) {
        case VR_INT32:
        case VR_UTF8: ((int32_t*)dst.data)[o] = ((const int32_t*)src.data)[i]; break;
        case VR_INT64:
        case VR_FLOAT64: ((int64_t*)dst.data)[o] = ((const int64_t*)src.data)[i]; break;
        case VR_DECIMAL128:
            ((int64_t*)dst.data)[2 * o] = ((const int64_t*)src.data)[2 * i];
            ((int64_t*)dst.data)[2 * o + 1] = ((const int64_t*)src.data)[2 * i + 1];
            break;
        case VR_BOOL:
            if (bit_at((const uint8_t*)src.data, i)) atomicOr((unsigned int*)dst.data + (o >> 5), 1u << (o & 31));
            break;
    }
}

// Writes output row o of a build column from a row's payload words staged in LDS (`row_sh`: word w
// at row_sh[w * PROBE_BLOCK + thread]; the lane starts at byte `off`); payload lanes have no nulls.
__device__ __forceinline__ void copy_pay(const uint32_t* row_sh, int off, const OutCol& dst, int64_t o) {
    const uint32_t* p = row_sh + (off / 4) * PROBE_BLOCK + threadIdx.x;
    if (dst.valid) atomicOr((unsigned int*)dst.valid + (o >> 5), 1u << (o & 31));
    switch (dst.type) {
        case VR_INT32: ((uint32_t*)dst.data)[o] = p[0]; break;
        case VR_INT64:
        case VR_FLOAT64: ((uint64_t*)dst.data)[o] = (uint64_t)p[0] | ((uint64_t)p[PROBE_BLOCK] << 32); break;
        case VR_DECIMAL128:
            ((uint64_t*)dst.data)[2 * o] = (uint64_t)p[0] | ((uint64_t)p[PROBE_BLOCK] << 32);
            ((uint64_t*)dst.data)[2 * o + 1] = (uint64_t)p[2 * PROBE_BLOCK] | ((uint64_t)p[3 * PROBE_BLOCK] << 32);
            break;
        default: break;
    }
}

// LDS bytes of a probe launch that reads the payload table: its rows' payload words.
__host__ __device__ inline size_t pay_lds(int words) { return (size_t)PROBE_ITEMS * words * PROBE_BLOCK * 4; }

constexpr uint64_t ST_AGG = 1ull << 62, ST_PREFIX = 2ull << 62, ST_MASK = 3ull << 62;
constexpr int PROBE_WAVES = PROBE_BLOCK / 64;
// Read in place of a selection or validity bitmap a batch does not have (every bit set).
__device__ const uint8_t k_probe_ones[8] = {0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF};

// Exclusive prefix sum of v over the wave's lanes; *total = the wave's sum.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// Decoupled look-back of tile `tile` (> 0) by one whole wave: each step reads the states of the 64
// tiles before the window's end at once, waits until every one of them up to the nearest inclusive
// prefix has published, and adds their values.  Returns the exclusive prefix of the tile.  Tile 0
// publishes an inclusive prefix without waiting, and every tile has its id before it waits, so
// every wait ends.
__device__ __forceinline__ uint64_t wave_look_back(uint64_t* tile_state, int64_t tile) {
    const int lane = threadIdx.x & 63;
    uint64_t excl = 0;
    int64_t end = tile - 1;  // the window is tiles end, end - 1, ..., end - 63
    while (true) {
        const int64_t idx = end - lane;
        const uint64_t st = idx >= 0 ? ld_coherent(&tile_state[idx]) : ST_PREFIX;
        const uint64_t status = st & ST_MASK;
        const uint64_t pmask = __ballot(status == ST_PREFIX);
        const uint64_t umask = __ballot(status == 0);
        const int limit = pmask ? __ffsll((unsigned long long)pmask) - 1 : 63;  // lanes 0..limit count
        const uint64_t need = limit == 63 ? ~0ull : ((2ull << limit) - 1);
        if (umask & need) continue;  // a tile in the window has not published yet
        uint64_t v = lane <= limit ? (st & ~ST_MASK) : 0;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
        excl += v;
        if (pmask) return excl;
        end -= 64;
    }
}

// The probe: tiles of PROBE_TILE streamed rows, taken in order from an atomic counter across every
// batch of the launch.  A tile's waves look their rows up (wave-cooperatively, the keys of all
// their rows loaded first), count each row's output rows and scan the counts within the wave, with
// no barrier between rounds; then the tile's offsets come from the waves' totals and its output
// offset from a wave-wide decoupled look-back over the preceding tiles, and the rows write pairs
// or materialized columns.  Outputs of all batches follow each other in batch order, each batch's
// in the reference's order: streamed rows in order, each row's build rows in build order.
template <bool MULTI, bool PAY>
__global__ __launch_bounds__(PROBE_BLOCK) void join_probe_kernel(ProbeArgs a) {
    __shared__ int64_t tile_id_sh;
    __shared__ int32_t batch_sh;
    __shared__ uint64_t base_sh;
    __shared__ uint32_t wave_tot[PROBE_ITEMS][PROBE_WAVES];  // totals, then their exclusive prefix in the tile
    __shared__ DCol psrc_sh[8], ksrc_sh[MAX_JOIN_KEYS];
    extern __shared__ uint32_t pay_sh[];  // payload table lookups: row r's word w at [(r * PAY_WORDS + w) * PROBE_BLOCK + thread]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        const int64_t t = (int64_t)atomic_add_u64(&a.tile_counter[0], 1);
        tile_id_sh = t;
        batch_sh = MULTI ? a.tile_batch[t] : 0;
    }
    __syncthreads();
    const int64_t tile = tile_id_sh;
    const int32_t bi = __builtin_amdgcn_readfirstlane(batch_sh);
    const PBatch B = MULTI ? a.batches[bi] : a.one;
    const PCol* bc = MULTI ? a.bcols + (int64_t)bi * a.ncb : nullptr;
    const int64_t row0 = (tile - B.tile0) * PROBE_TILE;
    const PCol k0 = slot_col<MULTI>(a, bc, 0);
    const PCol k1 = a.nk > 1 ? slot_col<MULTI>(a, bc, 1) : PCol{nullptr, nullptr};
    const void* pd[MAX_JOIN_KEYS] = {nullptr, nullptr, nullptr, nullptr};  // wide keys: the streamed key lanes
    const uint8_t* pv[MAX_JOIN_KEYS] = {nullptr, nullptr, nullptr, nullptr};
    if (a.t.wide)
        for (int c = 0; c < a.nk; c++) {
            const PCol p = slot_col<MULTI>(a, bc, c);
            pd[c] = p.data;
            pv[c] = p.valid;
        }
    DCol resl = a.res_lbuild, resr = a.res_rbuild;
    if (a.res_on) {
        if (!a.res_lb) {
            const PCol p = slot_col<MULTI>(a, bc, a.res_lslot);
            resl = DCol{p.valid, p.data, nullptr, a.ctype[a.res_lslot], 0};
        }
        if (!a.res_rb) {
            const PCol p = slot_col<MULTI>(a, bc, a.res_rslot);
            resr = DCol{p.valid, p.data, nullptr, a.ctype[a.res_rslot], 0};
        }
    }
    // the keys of all the thread's rows first: their loads overlap
    uint64_t key[PROBE_ITEMS];
    uint32_t flags = 0;  // bit r: row r is in (selected), bit 16 + r: its key is non-null
    if (a.t.wide) {
#pragma unroll
        for (int r = 0; r < PROBE_ITEMS; r++) {
            const int64_t i = row0 + (int64_t)r * PROBE_BLOCK + threadIdx.x;
            const bool in = i < B.n && (!B.sel || bit_at(B.sel, i));
            key[r] = 0;
            bool nonnull = in;  // the key's hash; a null lane makes the key null
            for (int c = 0; c < a.nk; c++)
                if (pv[c] && !bit_at(pv[c], i)) nonnull = false;
            if (nonnull) key[r] = wide_hash(pd, a.bk.width, a.nk, a.bk.seed, i);
            flags |= (in ? 1u : 0u) << r;
            flags |= (nonnull ? 1u : 0u) << (16 + r);
        }
    } else {
        // One INT32/INT64/FLOAT64 key or two INT32 keys (probe_key's cases): every load of the
        // rows issued before any is used, with no branch between them -- the row clamped into the
        // batch, bitmaps the batch lacks read from a constant byte (offset masked to 0), the second
        // word of a one-lane key read from its own lane -- so the tile waits for one round trip,
        // not for a load per bitmap and row.
        const int64_t n = B.n;  // >= 1: a tile holds rows of its batch
        const bool two = a.nk > 1;
        const bool w8 = !two && key_lane_width(a.ctype[0]) == 8;
        const uint8_t* sp = B.sel ? B.sel : k_probe_ones;
        const int64_t sm = B.sel ? -1 : 0;
        const uint8_t* v0 = k0.valid ? k0.valid : k_probe_ones;
        const int64_t m0 = k0.valid ? -1 : 0;
        const uint8_t* v1 = two && k1.valid ? k1.valid : k_probe_ones;
        const int64_t m1 = two && k1.valid ? -1 : 0;
        const uint32_t* d0 = (const uint32_t*)k0.data;
        const uint32_t* d1 = two ? (const uint32_t*)k1.data : d0;
        uint32_t sb[PROBE_ITEMS], b0[PROBE_ITEMS], b1[PROBE_ITEMS], lo[PROBE_ITEMS], hi[PROBE_ITEMS];
#pragma unroll
        for (int r = 0; r < PROBE_ITEMS; r++) {
            const int64_t i = row0 + (int64_t)r * PROBE_BLOCK + threadIdx.x;
            const int64_t ci = i < n ? i : n - 1;
            sb[r] = *gptr(sp + ((ci >> 3) & sm));
            b0[r] = *gptr(v0 + ((ci >> 3) & m0));
            b1[r] = *gptr(v1 + ((ci >> 3) & m1));
            const int64_t e0 = w8 ? 2 * ci : ci;
            lo[r] = ld_stream(d0 + e0);
            hi[r] = ld_stream(w8 ? d0 + e0 + 1 : d1 + ci);
        }
#pragma unroll
        for (int r = 0; r < PROBE_ITEMS; r++) {
            const int64_t i = row0 + (int64_t)r * PROBE_BLOCK + threadIdx.x;
            const int bit = (int)(i & 7);
            // bitwise, not short-circuit: a load whose value is needed only behind a branch would
            // be sunk into it (and waited for there)
            const uint32_t in_bits = (uint32_t)(i < n) & (sb[r] >> bit);
            const bool in = in_bits & 1u;
            const bool nonnull = (in_bits & (b0[r] >> bit) & (b1[r] >> bit)) & 1u;
            const uint64_t k = two ? (((uint64_t)lo[r] << 32) | hi[r]) : (w8 ? ((uint64_t)lo[r] | ((uint64_t)hi[r] << 32)) : lo[r]);
            key[r] = nonnull ? k : 0;
            flags |= (in ? 1u : 0u) << r;
            flags |= (nonnull ? 1u : 0u) << (16 + r);
        }
    }
    uint32_t cnt[PROBE_ITEMS], off[PROBE_ITEMS];
    uint32_t payload[PROBE_ITEMS], found[PROBE_ITEMS], npass[PROBE_ITEMS];
    const bool pairs = a.join_type == VR_JOIN_INNER || a.join_type == VR_JOIN_OUTER;
#pragma unroll
    for (int r = 0; r < PROBE_ITEMS; r++) {
        const int64_t i = row0 + (int64_t)r * PROBE_BLOCK + threadIdx.x;
        const bool in = (flags >> r) & 1u, nonnull = (flags >> (16 + r)) & 1u;
        uint64_t v = 0;
        if (a.build_rows) {  // uniform: every lane of the wave takes the same branch
            if (PAY) {  // the payload table (a.pt)
                uint32_t pw[PAY_WORDS];
                const int64_t br = coop_pay_lookup(a.pt, nonnull && key[r] != EMPTY_KEY, key[r], pw);
                v = br >= 0 ? ((1ull << 32) | (uint64_t)br) : 0ull;
#pragma unroll
                for (int w = 0; w < PAY_WORDS; w++)
                    if (w < a.pay_words) pay_sh[(r * a.pay_words + w) * PROBE_BLOCK + threadIdx.x] = pw[w];
            } else if (a.dense.slots) {
                if (nonnull) v = dense_lookup(a.dense, a.ctype[0], key[r]);
            } else {
                v = coop_lookup(a.t, nonnull && key[r] != EMPTY_KEY, key[r]);
                if (nonnull && key[r] == EMPTY_KEY) v = a.t.sentinel_value;
                if (a.t.wide && v) {  // the run's key must be the row's: its first build row decides
                    const uint32_t m = (uint32_t)(v >> 32), pl = (uint32_t)v;
                    const int64_t first = m == 1 ? (int64_t)pl : (int64_t)a.t.range_rows[pl];
                    if (!wide_equal(pd, i, a.bk.data, first, a.bk.width, a.nk)) v = 0;
                }
            }
        }
        uint32_t c = 0, pl = 0, fd = 0, np = 0;
        if (in) {
            const uint32_t matches = (uint32_t)(v >> 32);
            pl = (uint32_t)v;
            fd = matches;
            np = matches;
            if (a.res_on && matches) {  // the candidates that pass the residual
                np = 0;
                for (uint32_t k = 0; k < matches; k++) np += res_passes(a, resl, resr, i, candidate(a, matches, pl, k)) ? 1u : 0u;
            }
            if (pairs) {
                c = np ? np : (a.join_type == VR_JOIN_OUTER ? 1u : 0u);  // (row, -1) when preserved
            } else {
                const bool matched = np > 0 || (a.drop_null_keys && a.build_rows && !nonnull);
                c = (matched == (a.join_type == VR_JOIN_LEFT_SEMI)) ? 1u : 0u;
            }
        }
        cnt[r] = c;
        payload[r] = pl;
        found[r] = fd;
        npass[r] = np;
        uint32_t total;
        off[r] = wave_excl_scan(c, &total);
        if (lane == 0) wave_tot[r][wave] = total;
    }
    __syncthreads();
    if (wave == 0) {
        // the waves' offsets within the tile, in row order: round-major, then wave
        uint64_t local = 0;
        if (lane == 0) {
            uint32_t run = 0;
            for (int r = 0; r < PROBE_ITEMS; r++)
                for (int w = 0; w < PROBE_WAVES; w++) {
                    const uint32_t t = wave_tot[r][w];
                    wave_tot[r][w] = run;
                    run += t;
                }
            local = run;
        }
        local = __shfl(local, 0);
        uint64_t excl = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&a.tile_state[0], ST_PREFIX | local, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&a.tile_state[tile], ST_AGG | local, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            excl = wave_look_back(a.tile_state, tile);
            if (lane == 0)
                __hip_atomic_store(&a.tile_state[tile], ST_PREFIX | (excl + local), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            if (tile == a.ntiles - 1) __hip_atomic_store(&a.tile_counter[1], excl + local, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (MULTI && a.batch_end && tile == B.tile0 + (B.n + PROBE_TILE - 1) / PROBE_TILE - 1)
                a.batch_end[bi] = (int64_t)(excl + local);
            base_sh = excl;
        }
    }
    // the streamed columns the rows write from, staged in LDS: read from the batch table inside
    // the loop, a descriptor would be read again after every store to an output (which could alias
    // the table), a memory round trip per column and row
    if (threadIdx.x < (unsigned)a.nprobe_out) {
        const PCol p = slot_col<MULTI>(a, bc, a.nk + threadIdx.x);
        psrc_sh[threadIdx.x] = DCol{p.valid, p.data, nullptr, a.ctype[a.nk + threadIdx.x], 0};
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + (unsigned)a.nk) {
        const int q = threadIdx.x - 64;
        const PCol p = slot_col<MULTI>(a, bc, q);
        ksrc_sh[q] = DCol{p.valid, p.data, nullptr, a.ctype[q], 0};
    }
    __syncthreads();
    const uint64_t base = base_sh;
    for (int r = 0; r < PROBE_ITEMS; r++) {
        const int64_t i = row0 + (int64_t)r * PROBE_BLOCK + threadIdx.x;
        uint64_t o = base + wave_tot[r][wave] + off[r];
        uint32_t k = 0;  // the next candidate of the row's key
        for (uint32_t m = 0; m < cnt[r]; m++, o++) {
            if ((int64_t)o >= a.capacity) break;
            int64_t brow = -1;
            if (pairs && npass[r]) {  // the next passing candidate; a row with none is padded
                while (k < found[r]) {
                    brow = candidate(a, found[r], payload[r], k++);
                    if (!a.res_on || res_passes(a, resl, resr, i, brow)) break;
                }
            }
            if (a.matched && brow >= 0) {  // buildMatched(r) = true; bits only ever get set
                uint32_t* w = a.matched + (brow >> 5);
                const uint32_t bit = 1u << (brow & 31);
                if (!(*w & bit)) atomicOr(w, bit);
            }
            if (a.probe_idx) a.probe_idx[o] = (int32_t)i;
            if (a.build_idx) a.build_idx[o] = (int32_t)brow;
            for (int c = 0; c < a.nprobe_out; c++) copy_value(psrc_sh[c], i, a.outs[c], (int64_t)o);
            for (int c = 0; c < a.nbuild_out; c++) {
                const int ks = a.build_from_key[c];
                if (ks >= 0 && brow >= 0) copy_value(ksrc_sh[ks], i, a.outs[a.nprobe_out + c], (int64_t)o);
                else if (PAY && a.pay_off[c] >= 0 && brow >= 0)
                    copy_pay(pay_sh + r * a.pay_words * PROBE_BLOCK, a.pay_off[c], a.outs[a.nprobe_out + c], (int64_t)o);
                else copy_value(a.build_src[c], brow, a.outs[a.nprobe_out + c], (int64_t)o);
            }
        }
    }
}

// The batch of every tile of a many-batch launch.
__global__ void tile_batch_kernel(const PBatch* batches, int32_t nbatch, int32_t* tile_batch) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbatch; b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t0 = batches[b].tile0, nt = (batches[b].n + PROBE_TILE - 1) / PROBE_TILE;
        for (int64_t k = 0; k < nt; k++) tile_batch[t0 + k] = (int32_t)b;
    }
}

__global__ void zero_u64_kernel(uint64_t* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void fill_keys_kernel(uint64_t* buckets, uint64_t nbuckets) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbuckets * 16; i += (uint64_t)gridDim.x * blockDim.x)
        buckets[i] = (i % 2) == 0 ? EMPTY_KEY : 0ull;
}
// Build statistics: stats[1] += cells holding rows (distinct keys), stats[2] = max rows of a cell.
__global__ __launch_bounds__(256) void cell_stats_kernel(const uint32_t* cell_count, uint64_t ncells, uint64_t* stats) {
    typedef hipcub::BlockReduce<uint32_t, 256> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    uint32_t used = 0, mx = 0;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t k = cell_count[c];
        used += k != 0;
        mx = k > mx ? k : mx;
    }
    uint32_t u = Reduce(tmp).Sum(used);
    __syncthreads();
    uint32_t m = Reduce(tmp).Reduce(mx, hipcub::Max());
    if (threadIdx.x == 0) {
        atomicAdd((unsigned long long*)&stats[1], (unsigned long long)u);
        atomicMax((unsigned long long*)&stats[2], (unsigned long long)m);
    }
}

// stats[3] = smallest, stats[4] = largest key of the occupied cells (signed values).
__global__ void key_range_kernel(const uint64_t* buckets, uint64_t nbuckets, int32_t type, uint64_t* stats) {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nbuckets * BUCKET; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = buckets[2 * c];
        if (key == EMPTY_KEY) continue;
        const int64_t v = key_value(type, key);
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    if (lo <= hi) {
        atomic_min_i64(&stats[3], lo);
        atomic_max_i64(&stats[4], hi);
    }
}
__global__ void range_init_kernel(uint64_t* stats) {
    if (threadIdx.x == 0) {
        stats[3] = (uint64_t)INT64_MAX;
        stats[4] = (uint64_t)INT64_MIN;
    }
}
// slots[key - min] = the cell's value word, for every occupied cell.
__global__ void dense_fill_kernel(const uint64_t* buckets, uint64_t nbuckets, int32_t type, int64_t min, uint64_t* slots) {
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nbuckets * BUCKET; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = buckets[2 * c];
        if (key == EMPTY_KEY) continue;
        slots[(uint64_t)key_value(type, key) - (uint64_t)min] = buckets[2 * c + 1] & ~PASSED;
    }
}
__global__ void store_u64_kernel(uint64_t* p, uint64_t v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ void iota_kernel(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (uint32_t)i;
}
__global__ void sentinel_rows_kernel(const uint32_t* row_cell, int64_t n, uint32_t* out, uint64_t* count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (row_cell[i] == 0xFFFFFFFEu) out[atomic_add_u64(count, 1)] = (uint32_t)i;
}

// hipcub selection of the build rows without their matched bit.
struct UnmatchedRow {
    const uint32_t* matched;
    __host__ __device__ __forceinline__ bool operator()(const int32_t& r) const {
        return !((matched[r >> 5] >> (r & 31)) & 1u);
    }
};

struct UnmatchedArgs {
    const int32_t* rows;  // the unmatched build rows, in build order
    int64_t m;
    int32_t* probe_idx;
    int32_t* build_idx;
    int32_t nprobe_out, nbuild_out;
    DCol build_src[8];
    OutCol outs[16];
};

// emitUnmatchedBuild's rows (VectorHashJoinExec.scala:1457-1484): (-1, r), the streamed columns
// null, the build columns gathered.
__global__ void unmatched_rows_kernel(UnmatchedArgs a) {
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < a.m; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = a.rows[o];
        if (a.probe_idx) a.probe_idx[o] = -1;
        if (a.build_idx) a.build_idx[o] = (int32_t)r;
        for (int c = 0; c < a.nprobe_out; c++) copy_value(DCol{}, -1, a.outs[c], o);
        for (int c = 0; c < a.nbuild_out; c++) copy_value(a.build_src[c], r, a.outs[a.nprobe_out + c], o);
    }
}

static int grid_of(uint64_t work, int block, int cap) {
    uint64_t g = (work + block - 1) / block;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)cap));
}

struct JoinImpl {
    Device* dev = nullptr;
    std::vector<int32_t> key_types, col_types, key_cols;
    // build columns, device resident, grown by appends
    struct BuildCol {
        DevBuf data, valid;
        bool has_valid = false;
        size_t cap_rows = 0;
    };
    std::vector<BuildCol> cols;
    int64_t rows = 0;
    bool finished = false;
    // table
    DevBuf buckets, range_rows, row_cell, cell_count, cell_start, sorted_cells, rows_in, cub_tmp, small;
    uint64_t nbuckets = 0;
    uint64_t sentinel_value = 0;
    int64_t distinct = 0, max_chain = 0;
    // build rows paired by probes (buildMatched), when the spec asks for them
    bool track = false;
    DevBuf matched, unmatched;
    // the payload table (unique keys, fixed-width payload lanes without nulls, at most 20 bytes):
    // pay_off[c] = build column c's byte offset in a cell's payload, -1 when not in it
    DevBuf pay;
    uint64_t pay_nb = 0;
    int32_t pay_words = 0;
    std::vector<int32_t> pay_off;
    // the direct key index (DenseKeyIndex), when the build allows it
    int32_t dense_mode = 0;  // vr_join_spec.dense_index
    DevBuf dense;
    int64_t dense_min = 0;
    uint64_t dense_range = 0;
    // Pinned staging of host-resident streamed columns, shared by the probes of the handle: held
    // under `stage_mu` from the copy until the kernel reading it has finished.  Every other piece of
    // a probe's scratch (tile states, output staging, batch tables) belongs to the call, so probes
    // of a finished table may run from several threads at once.
    std::mutex stage_mu;
    Staging staging;
    // wide keys (more than 64 bits of key lanes): cells hold fingerprints and build rows
    bool wide = false;
    uint64_t seed = 0x9E3779B97F4A7C15ull;  // wide keys: the hash's seed (another on a collision)
    // UTF8 keys: each one's dictionary of the build side's strings (a build column of a UTF8 key
    // holds the rows' ids; a probe maps its strings to those ids, unknown ones to -1)
    std::vector<StrTable> strs;
};

__global__ void remap_codes_kernel(const int32_t* codes, int64_t n, const int32_t* remap, int32_t m, int32_t* ids) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = codes[i];
        ids[i] = c >= 0 && c < m ? remap[c] : -1;
    }
}

// ids[0, n) (device) of UTF8 key k's column `col`: inserting the build side's strings, or looking
// a probe's up (unknown: -1).  A dictionary-encoded column maps its dictionary's entries, then its
// codes through them.
static vr_status map_key_strings(JoinImpl& j, int k, const vr_column& col, int64_t n, bool insert, int32_t* ids,
                                 hipStream_t s) {
    StrTable& t = j.strs[(size_t)k];
    if (col.dictionary) {
        const vr_column& dc = *col.dictionary;
        if (dc.mem != VR_MEM_HOST || dc.type != VR_UTF8 || !dc.offsets)
            return fail(VR_ERR_UNSUPPORTED, "a key dictionary must be a plain UTF8 column in host memory");
        if (dc.validity)
            for (int64_t e = 0; e < dc.length; e++)
                if (!((dc.validity[e >> 3] >> (e & 7)) & 1))
                    return fail(VR_ERR_UNSUPPORTED, "join keys over a dictionary with null entries run on the CPU kernels");
        const int32_t m = (int32_t)dc.length;
        DevBuf remap, codes;
        VR_TRY(remap.ensure(j.dev, (size_t)std::max(m, 1) * 4));
        VR_TRY(t.map_host(j.dev, dc.offsets, (const uint8_t*)dc.data, nullptr, m, insert, -1, remap.as<int32_t>(), s));
        const int32_t* cp = (const int32_t*)col.data;
        if (col.mem == VR_MEM_HOST) {
            VR_TRY(codes.ensure(j.dev, (size_t)n * 4));
            VR_TRY_HIP(hipMemcpyAsync(codes.p, col.data, (size_t)n * 4, hipMemcpyHostToDevice, s));
            cp = codes.as<int32_t>();
        }
        hipLaunchKernelGGL(remap_codes_kernel, dim3(grid_of(n, 256, 4096)), dim3(256), 0, s, cp, n, remap.as<int32_t>(), m, ids);
        VR_TRY_HIP(hipGetLastError());
        VR_TRY_HIP(hipStreamSynchronize(s));  // the map and codes are freed on return
        return VR_OK;
    }
    if (!col.offsets) return fail(VR_ERR_INVALID_ARGUMENT, "a plain UTF8 key column needs its offsets");
    if (col.mem == VR_MEM_HOST) return t.map_host(j.dev, col.offsets, (const uint8_t*)col.data, col.validity, n, insert, 0, ids, s);
    return t.map(j.dev, StrRows{col.offsets, (const uint8_t*)col.data, col.validity, n}, insert, 0, ids, s);
}

static vr_status grow_col(JoinImpl& j, JoinImpl::BuildCol& c, int32_t type, size_t want_rows, hipStream_t s) {
    if (want_rows <= c.cap_rows) return VR_OK;
    size_t rows = std::max<size_t>(want_rows, c.cap_rows * 2);
    rows = std::max<size_t>(rows, 1024);
    DevBuf nd, nv;
    VR_TRY(nd.ensure(j.dev, lane_bytes(type, (int64_t)rows) + 8));
    VR_TRY(nv.ensure(j.dev, align_up(bitmap_bytes((int64_t)rows), 8) + 8));
    VR_TRY_HIP(hipMemsetAsync(nv.p, 0xFF, nv.bytes, s));
    if (j.rows) {
        VR_TRY_HIP(hipMemcpyAsync(nd.p, c.data.p, lane_bytes(type, j.rows), hipMemcpyDeviceToDevice, s));
        VR_TRY_HIP(hipMemcpyAsync(nv.p, c.valid.p, bitmap_bytes(j.rows), hipMemcpyDeviceToDevice, s));
    }
    VR_TRY_HIP(hipStreamSynchronize(s));
    c.data = std::move(nd);
    c.valid = std::move(nv);
    c.cap_rows = rows;
    return VR_OK;
}

__global__ void append_bits_kernel(uint8_t* dst, int64_t at, const uint8_t* src, int64_t n) {
    // dst bits [at, at + n) = src bits [0, n) (src NULL: all valid)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        bool v = !src || bit_at(src, i);
        int64_t o = at + i;
        unsigned int* w = (unsigned int*)dst + (o >> 5);
        if (v) atomicOr(w, 1u << (o & 31));
        else atomicAnd(w, ~(1u << (o & 31)));
    }
}

// The build side's key lanes, for wide-key inserts and comparisons.
static WideKeys wide_keys(const JoinImpl& j) {
    WideKeys wk{};
    wk.nk = (int32_t)j.key_types.size();
    for (int q = 0; q < wk.nk; q++) {
        const auto& bc = j.cols[j.key_cols[q]];
        wk.data[q] = bc.data.p;
        wk.valid[q] = bc.has_valid ? bc.valid.as<uint8_t>() : nullptr;
        wk.width[q] = key_lane_width(j.key_types[q]);
    }
    wk.seed = j.seed;
    return wk;
}

}  // namespace vr

using namespace vr;

struct vr_join {
    DeviceRef owner;  // destroyed after the state below
    JoinImpl j;
};

// The largest key range indexed directly, relative to the distinct keys and in slots
// (DenseKeyIndex.DENSE_FACTOR, MAX_RANGE: DenseKeyIndex.java:36-43).
constexpr int64_t DENSE_FACTOR = 10;
constexpr uint64_t DENSE_MAX_RANGE = 1ull << 22;

// DenseKeyIndex.tryBuild (DenseKeyIndex.java:70-110) over the finished table: the key range of the
// non-null build keys [lo, hi] (the empty marker's key, held beside the table, included), and when
// it spans at most DENSE_MAX_RANGE slots and DENSE_FACTOR times the distinct keys, one value word
// per key of the range.
static vr_status build_dense(JoinImpl& j, bool candidate, int64_t lo, int64_t hi, bool marker, hipStream_t s) {
    j.dense.reset();
    j.dense_range = 0;
    if (!candidate) return VR_OK;
    if (marker && j.key_types[0] == VR_INT64) {
        lo = std::min(lo, (int64_t)EMPTY_KEY);
        hi = std::max(hi, (int64_t)EMPTY_KEY);
    }
    if (lo > hi) return VR_OK;
    const uint64_t span = (uint64_t)hi - (uint64_t)lo;  // exact: hi >= lo
    if (span >= DENSE_MAX_RANGE || span + 1 > (uint64_t)(DENSE_FACTOR * j.distinct)) return VR_OK;
    const uint64_t range = span + 1;
    VR_TRY(j.dense.ensure(j.dev, range * sizeof(uint64_t)));
    const int cap = j.dev->num_cus * 16;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(range, 256, cap)), dim3(256), 0, s, j.dense.as<uint64_t>(), range);
    hipLaunchKernelGGL(dense_fill_kernel, dim3(grid_of(j.nbuckets * BUCKET, 256, cap)), dim3(256), 0, s, j.buckets.as<uint64_t>(),
                       j.nbuckets, j.key_types[0], lo, j.dense.as<uint64_t>());
    if (marker && j.key_types[0] == VR_INT64)
        hipLaunchKernelGGL(store_u64_kernel, dim3(1), dim3(64), 0, s, j.dense.as<uint64_t>() + (EMPTY_KEY - (uint64_t)lo),
                           j.sentinel_value);
    VR_TRY_HIP(hipGetLastError());
    j.dense_min = lo;
    j.dense_range = range;
    j.dev->count(&vr_stats::dense_join_builds, 1);
    j.dev->count(&vr_stats::kernel_launches, marker ? 3 : 2);
    return VR_OK;
}

// The allocation-free probe's scratch: the tiles' look-back state and the output row count,
// in a device array of the library (each device holds its own copy from when the library's code
// is loaded on it until the process ends).  Host code never passes the array's name to a kernel
// or a copy: the name is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress, once per device.
constexpr uint64_t NOALLOC_SCRATCH_WORDS = 1ull << 17;  // 1 MiB: two counters + 131,070 tiles
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_WORDS];

static std::mutex g_scratch_mu;
static std::unordered_map<int, uint64_t*> g_scratch_ptr;

// The address of `scratch` on device `ordinal`, which must be the current device.
static vr_status scratch_on(int ordinal, uint64_t** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(ordinal);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        VR_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(VR_ERR_DEVICE, "no device address for the probe scratch");
        it = g_scratch_ptr.emplace(ordinal, (uint64_t*)p).first;
    }
    *out = it->second;
    return VR_OK;
}

extern "C" {

vr_status vr_join_create(vr_device* dev, const vr_join_spec* spec, vr_join** out) {
    if (!dev || !spec || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (spec->num_keys < 1 || spec->num_keys > MAX_JOIN_KEYS) return fail(VR_ERR_UNSUPPORTED, "one to four join keys run on the device");
    if (spec->num_columns < 1 || spec->num_columns > 8) return fail(VR_ERR_UNSUPPORTED, "at most 8 build columns");
    std::unique_ptr<vr_join> h(new vr_join());
    h->owner.hold(dev);
    JoinImpl& j = h->j;
    j.dev = &dev->d;
    j.key_types.assign(spec->key_types, spec->key_types + spec->num_keys);
    j.col_types.assign(spec->column_types, spec->column_types + spec->num_columns);
    j.key_cols.assign(spec->key_columns, spec->key_columns + spec->num_keys);
    int key_bytes = 0;
    for (int k = 0; k < spec->num_keys; k++) {
        int32_t t = j.key_types[k];
        if (j.key_cols[k] < 0 || j.key_cols[k] >= spec->num_columns || j.col_types[j.key_cols[k]] != t)
            return fail(VR_ERR_INVALID_ARGUMENT, "join key column does not match the build columns");
        if (t == VR_BOOL) return fail(VR_ERR_UNSUPPORTED, "BOOL join keys run on the CPU kernels");
        key_bytes += key_lane_width(t);
    }
    // one lane of at most 8 bytes, or two of 4: the key is its 64 bits; else cells hold fingerprints
    j.wide = !(spec->num_keys == 1 && key_bytes <= 8) && !(spec->num_keys == 2 && key_bytes == 8);
    for (int c = 0; c < spec->num_columns; c++) {
        if (j.col_types[c] != VR_UTF8) continue;
        bool key = false;
        for (int k = 0; k < spec->num_keys; k++) key |= j.key_cols[k] == c;
        if (!key) return fail(VR_ERR_UNSUPPORTED, "UTF8 payloads run on the CPU kernels");
    }
    j.strs.resize((size_t)spec->num_keys);
    if (spec->dense_index != 0 && spec->dense_index != -1) return fail(VR_ERR_INVALID_ARGUMENT, "dense_index is 0 or -1");
    j.dense_mode = spec->dense_index;
    j.track = spec->track_matched != 0;
    j.cols.resize(spec->num_columns);
    VR_TRY_HIP(hipSetDevice(j.dev->ordinal));
    if (spec->capacity_hint > 0)
        for (int c = 0; c < spec->num_columns; c++)
            VR_TRY(grow_col(j, j.cols[c], j.col_types[c], (size_t)spec->capacity_hint, j.dev->stream));
    *out = h.release();
    return VR_OK;
}

vr_status vr_join_build_append(vr_join* h, const vr_column* columns, int64_t n, void* stream) {
    if (!h || (!columns && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    if (j.finished) return fail(VR_ERR_STATE, "the build side is already finished");
    if (n <= 0) return n < 0 ? fail(VR_ERR_INVALID_ARGUMENT, "negative row count") : VR_OK;
    if (j.rows + n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "build sides past 2^31 rows run on the CPU kernels");
    hipStream_t s = j.dev->pick(stream);
    for (size_t c = 0; c < j.cols.size(); c++) {
        const vr_column& col = columns[c];
        if (col.type != j.col_types[c]) return fail(VR_ERR_INVALID_ARGUMENT, "build column lane type differs from the spec");
        if (col.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "build column shorter than the batch");
        auto& bc = j.cols[c];
        VR_TRY(grow_col(j, bc, col.type, (size_t)(j.rows + n), s));
        hipMemcpyKind kind = col.mem == VR_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (col.type == VR_BOOL) {
            return fail(VR_ERR_UNSUPPORTED, "BOOL build payloads run on the CPU kernels");
        }
        size_t w = lane_bytes(col.type, 1);
        if (col.type == VR_UTF8) {  // a UTF8 key: the rows' ids in the key's dictionary
            int k = 0;
            while (j.key_cols[k] != (int32_t)c) k++;
            VR_TRY(map_key_strings(j, k, col, n, true, bc.data.as<int32_t>() + j.rows, s));
        } else {
            VR_TRY_HIP(hipMemcpyAsync(bc.data.as<uint8_t>() + (size_t)j.rows * w, col.data, (size_t)n * w, kind, s));
        }
        if (col.validity) {
            const uint8_t* src = col.validity;
            DevBuf tmp;
            if (col.mem == VR_MEM_HOST) {
                VR_TRY(tmp.ensure(j.dev, bitmap_bytes(n)));
                VR_TRY_HIP(hipMemcpyAsync(tmp.p, col.validity, bitmap_bytes(n), hipMemcpyHostToDevice, s));
                src = tmp.as<uint8_t>();
            }
            hipLaunchKernelGGL(append_bits_kernel, dim3(grid_of(n, 256, 1024)), dim3(256), 0, s, bc.valid.as<uint8_t>(),
                               j.rows, src, n);
            VR_TRY_HIP(hipGetLastError());
            if (col.mem == VR_MEM_HOST) VR_TRY_HIP(hipStreamSynchronize(s));
            bc.has_valid = true;
        } else if (bc.has_valid) {
            hipLaunchKernelGGL(append_bits_kernel, dim3(grid_of(n, 256, 1024)), dim3(256), 0, s, bc.valid.as<uint8_t>(),
                               j.rows, (const uint8_t*)nullptr, n);
        }
        if (col.mem == VR_MEM_HOST) j.dev->count(&vr_stats::h2d_bytes, (int64_t)((size_t)n * w));
    }
    VR_TRY_HIP(hipStreamSynchronize(s));
    j.rows += n;
    j.dev->count(&vr_stats::input_rows, n);
    return VR_OK;
}

// The payload table, when every key is unique (no key equal to the empty marker) and the non-key
// build columns are fixed-width lanes without nulls that fit a cell's 20 payload bytes.
static vr_status build_payload_table(JoinImpl& j, bool marker, hipStream_t s) {
    j.pay.reset();
    j.pay_nb = 0;
    j.pay_off.assign(j.cols.size(), -1);
    const int64_t n = j.rows;
    if (j.wide || marker || n <= 0 || j.max_chain != 1 || n >= (int64_t)PROW_PASSED) return VR_OK;
    PaySrc src{};
    int off = 0;
    for (size_t c = 0; c < j.cols.size(); c++) {
        if (std::find(j.key_cols.begin(), j.key_cols.end(), (int32_t)c) != j.key_cols.end()) continue;
        const int32_t t = j.col_types[c];
        const int w = t == VR_INT32 ? 4 : (t == VR_INT64 || t == VR_FLOAT64) ? 8 : t == VR_DECIMAL128 ? 16 : 0;
        if (!w || j.cols[c].has_valid || off + w > PAY_WORDS * 4 || src.ncols == PAY_MAX_COLS) return VR_OK;
        src.data[src.ncols] = j.cols[c].data.p;
        src.width[src.ncols] = w;
        src.off[src.ncols] = off;
        src.ncols++;
        j.pay_off[c] = off;
        off += w;
    }
    if (!src.ncols) {
        j.pay_off.assign(j.cols.size(), -1);
        return VR_OK;
    }
    // a quarter of the cells hold keys: with four cells to a bucket, about one home bucket in 270
    // overflows (at half, one in 19), and a row whose home bucket overflowed holds its wave back
    const uint64_t cells = (uint64_t)((double)n / 0.25) + 1;
    const uint64_t nb = (cells + PBUCKET - 1) / PBUCKET;
    VR_TRY(j.pay.ensure(j.dev, nb * 16 * sizeof(uint64_t)));
    const int cap = j.dev->num_cus * 16;
    hipLaunchKernelGGL(pay_fill_kernel, dim3(grid_of(nb * 16, 256, cap)), dim3(256), 0, s, j.pay.as<uint64_t>(), nb * 16);
    DCol k[2]{};
    for (size_t q = 0; q < j.key_types.size(); q++) {
        auto& bc = j.cols[j.key_cols[q]];
        k[q].type = j.key_types[q];
        k[q].data = bc.data.p;
        k[q].valid = bc.has_valid ? bc.valid.as<uint8_t>() : nullptr;
    }
    hipLaunchKernelGGL(pay_insert_kernel, dim3(grid_of(n, 256, cap)), dim3(256), 0, s, PayTable{j.pay.as<uint64_t>(), nb}, k[0],
                       k[1], (int)j.key_types.size(), n, src);
    VR_TRY_HIP(hipGetLastError());
    j.pay_nb = nb;
    j.pay_words = (off + 3) / 4;
    return VR_OK;
}

vr_status vr_join_build_finish(vr_join* h, void* stream) {
    if (!h) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    if (j.finished) return fail(VR_ERR_STATE, "the build side is already finished");
    hipStream_t s = j.dev->pick(stream);
    const int64_t n = j.rows;
    // at most half of the cells hold keys.  A probe reads one bucket line per row unless its home
    // bucket was passed; a few such rows hold their whole wave back, so fewer of them pays more than
    // a table that stays in the 256 MB MALL: 10M build keys at 70% load (218 MB) probed 400M rows
    // in 15.4 ms (semi join), at 50% (305 MB) in 13.5 ms, at 40% in 13.3 ms
    const uint64_t cells_wanted = (uint64_t)((double)std::max<int64_t>(n, 1) / 0.50) + 1;
    const uint64_t nb = (cells_wanted + BUCKET - 1) / BUCKET;
    j.nbuckets = nb;
    const uint64_t ncells = nb * BUCKET;
    int cap = j.dev->num_cus * 16;
    const bool dense_candidate = j.dense_mode == 0 && j.key_types.size() == 1 && n > 0 &&
                                 (j.key_types[0] == VR_INT32 || j.key_types[0] == VR_INT64);
    DevBuf hashes, hvalid;  // wide keys: the build rows' hashes and which rows have no null key lane
    uint64_t small[5] = {0, 0, 0, 0, 0};
    uint64_t sentinel_rows = 0;
    for (int attempt = 0;; attempt++) {
        VR_TRY(j.buckets.ensure(j.dev, nb * 16 * sizeof(uint64_t)));
        hipLaunchKernelGGL(fill_keys_kernel, dim3(grid_of(nb * 16, 256, cap)), dim3(256), 0, s, j.buckets.as<uint64_t>(), nb);
        VR_TRY(j.row_cell.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 4));
        VR_TRY(j.cell_count.ensure(j.dev, ncells * 4 + 8));
        VR_TRY(j.cell_start.ensure(j.dev, ncells * 4 + 8));
        VR_TRY(j.small.ensure(j.dev, 64));
        const uint64_t passed_words = (nb + 63) / 64;  // one bit per bucket, in 64-bit words
        DevBuf passed;
        VR_TRY(passed.ensure(j.dev, passed_words * 8));
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(ncells / 2 + 1, 256, cap)), dim3(256), 0, s, j.cell_count.as<uint64_t>(),
                           ncells / 2 + 1);
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(passed_words, 256, cap)), dim3(256), 0, s, passed.as<uint64_t>(), passed_words);
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, j.small.as<uint64_t>(), 8ull);
        JoinTable t{j.buckets.as<uint64_t>(), nb, nullptr, 0, j.wide ? 1 : 0, 0};
        if (j.wide) {  // the rows' hashes, inserted as 64-bit keys
            VR_TRY(hashes.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 8));
            VR_TRY(hvalid.ensure(j.dev, ((size_t)(n + 31) / 32 + 2) * 4));
            const WideKeys wk = wide_keys(j);
            if (n > 0) {
                hipLaunchKernelGGL(wide_hash_kernel, dim3(grid_of((n + 31) / 32, 256, cap)), dim3(256), 0, s, wk, n,
                                   hashes.as<uint64_t>(), hvalid.as<uint32_t>());
                const DCol hk{hvalid.as<uint8_t>(), hashes.p, nullptr, VR_INT64, 0};
                hipLaunchKernelGGL(join_insert_kernel, dim3(grid_of(n, 256, cap)), dim3(256), 0, s, t, hk, DCol{}, 1, n,
                                   j.row_cell.as<uint32_t>(), j.cell_count.as<uint32_t>(), j.small.as<uint64_t>(),
                                   passed.as<uint32_t>());
            }
        } else {
            DCol k[2]{};
            for (size_t q = 0; q < j.key_types.size(); q++) {
                auto& bc = j.cols[j.key_cols[q]];
                k[q].type = j.key_types[q];
                k[q].data = bc.data.p;
                k[q].valid = bc.has_valid ? bc.valid.as<uint8_t>() : nullptr;
            }
            if (n > 0)
                hipLaunchKernelGGL(join_insert_kernel, dim3(grid_of(n, 256, cap)), dim3(256), 0, s, t, k[0], k[1],
                                   (int)j.key_types.size(), n, j.row_cell.as<uint32_t>(), j.cell_count.as<uint32_t>(),
                                   j.small.as<uint64_t>(), passed.as<uint32_t>());
        }
        VR_TRY_HIP(hipGetLastError());
        // run starts, then the build rows ordered by cell (stable: build order inside a key)
        size_t tmp = 0;
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, j.cell_count.as<uint32_t>(), j.cell_start.as<uint32_t>(),
                                                    (int64_t)ncells, s));
        VR_TRY(j.cub_tmp.ensure(j.dev, tmp));
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(j.cub_tmp.p, tmp, j.cell_count.as<uint32_t>(), j.cell_start.as<uint32_t>(),
                                                    (int64_t)ncells, s));
        VR_TRY(j.range_rows.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 4));
        VR_TRY(j.sorted_cells.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 4));
        VR_TRY(j.rows_in.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 4));
        if (n > 0) {
            hipLaunchKernelGGL(iota_kernel, dim3(grid_of(n, 256, cap)), dim3(256), 0, s, j.rows_in.as<uint32_t>(), n);
            size_t t2 = 0;
            VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, j.row_cell.as<uint32_t>(), j.sorted_cells.as<uint32_t>(),
                                                          j.rows_in.as<uint32_t>(), j.range_rows.as<uint32_t>(), (int)n, 0, 32, s));
            VR_TRY(j.cub_tmp.ensure(j.dev, std::max(tmp, t2)));
            VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(j.cub_tmp.p, t2, j.row_cell.as<uint32_t>(), j.sorted_cells.as<uint32_t>(),
                                                          j.rows_in.as<uint32_t>(), j.range_rows.as<uint32_t>(), (int)n, 0, 32, s));
        }
        hipLaunchKernelGGL(join_values_kernel, dim3(grid_of(ncells, 256, cap)), dim3(256), 0, s, t, j.cell_count.as<uint32_t>(),
                           j.cell_start.as<uint32_t>(), ncells, j.range_rows.as<uint32_t>(), passed.as<uint32_t>());
        VR_TRY_HIP(hipGetLastError());
        // statistics (distinct keys, longest chain) reduced on the device into small[1], small[2]
        hipLaunchKernelGGL(cell_stats_kernel, dim3(grid_of(ncells, 256, cap)), dim3(256), 0, s, j.cell_count.as<uint32_t>(),
                           ncells, j.small.as<uint64_t>());
        VR_TRY_HIP(hipGetLastError());
        if (dense_candidate) {
            hipLaunchKernelGGL(range_init_kernel, dim3(1), dim3(64), 0, s, j.small.as<uint64_t>());
            hipLaunchKernelGGL(key_range_kernel, dim3(grid_of(ncells, 256, cap)), dim3(256), 0, s, j.buckets.as<uint64_t>(), nb,
                               j.key_types[0], j.small.as<uint64_t>());
            VR_TRY_HIP(hipGetLastError());
        }
        // the key equal to the empty marker: its rows follow every cell's rows in range_rows
        uint32_t last[2] = {0, 0};
        VR_TRY_HIP(hipMemcpyAsync(small, j.small.p, sizeof(small), hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipMemcpyAsync(&last[0], j.cell_start.as<uint32_t>() + ncells - 1, 4, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipMemcpyAsync(&last[1], j.cell_count.as<uint32_t>() + ncells - 1, 4, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipStreamSynchronize(s));
        sentinel_rows = small[0];
        j.sentinel_value = 0;
        if (sentinel_rows) {
            // sorted by cell, the 0xFFFFFFFE rows sit just before the 0xFFFFFFFF (null) rows, in build order
            uint32_t total_cells_rows = last[0] + last[1];
            uint64_t payload = total_cells_rows;
            if (sentinel_rows == 1) {
                uint32_t row = 0;
                VR_TRY_HIP(hipMemcpyAsync(&row, j.range_rows.as<uint32_t>() + total_cells_rows, 4, hipMemcpyDeviceToHost, s));
                VR_TRY_HIP(hipStreamSynchronize(s));
                payload = row;
            }
            j.sentinel_value = (sentinel_rows << 32) | payload;
        }
        j.distinct = (int64_t)small[1] + (sentinel_rows ? 1 : 0);
        j.max_chain = std::max<int64_t>((int64_t)small[2], (int64_t)sentinel_rows);
        if (!j.wide || n == 0) break;
        // wide keys: every run must hold one key (two keys with one 64-bit hash share a cell)
        uint64_t collisions = 0;
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, j.small.as<uint64_t>() + 6, 1ull);
        hipLaunchKernelGGL(wide_verify_kernel, dim3(grid_of(n, 256, cap)), dim3(256), 0, s, wide_keys(j), n,
                           j.row_cell.as<uint32_t>(), j.cell_start.as<uint32_t>(), j.range_rows.as<uint32_t>(),
                           last[0] + last[1], j.small.as<uint64_t>() + 6);
        VR_TRY_HIP(hipGetLastError());
        VR_TRY_HIP(hipMemcpyAsync(&collisions, j.small.as<uint64_t>() + 6, 8, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipStreamSynchronize(s));
        if (!collisions) break;
        if (attempt == 3) return fail(VR_ERR_UNSUPPORTED, "wide join keys whose hashes keep colliding run on the CPU kernels");
        j.seed = fmix64(j.seed + (uint64_t)attempt + 1);  // build again with another hash
    }
    VR_TRY(build_dense(j, dense_candidate, (int64_t)small[3], (int64_t)small[4], sentinel_rows != 0, s));
    VR_TRY(build_payload_table(j, sentinel_rows != 0, s));
    if (j.track) {
        const uint64_t words = ((uint64_t)std::max<int64_t>(n, 1) + 63) / 64;
        VR_TRY(j.matched.ensure(j.dev, words * 8));
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(words, 256, cap)), dim3(256), 0, s, j.matched.as<uint64_t>(), words);
        VR_TRY_HIP(hipGetLastError());
    }
    j.row_cell.reset();
    j.sorted_cells.reset();
    j.rows_in.reset();
    j.cell_start.reset();
    j.cell_count.reset();
    j.finished = true;
    j.dev->count(&vr_stats::kernel_launches, 7);
    return VR_OK;
}

vr_status vr_join_info(vr_join* h, int64_t* build_rows, int64_t* distinct_keys, int64_t* max_rows_per_key) {
    if (!h) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (build_rows) *build_rows = h->j.rows;
    if (distinct_keys) *distinct_keys = h->j.distinct;
    if (max_rows_per_key) *max_rows_per_key = h->j.max_chain;
    return VR_OK;
}

// Where a call's outputs go: the caller's device buffers, or a device staging area of the call's
// own, copied back to the caller's host buffers after the kernel.
struct OutPlan {
    bool host = false;
    DevBuf stage;
    uint8_t* ob = nullptr;
    size_t pidx_off = 0, bidx_off = 0;
    std::vector<size_t> voff, doff;
};

// Points the kernel's pair arrays and output columns (of `types`) at their buffers and clears the
// validity words (bits are OR-ed in).
static vr_status plan_outputs(JoinImpl& j, const vr_probe_output* out, const std::vector<int32_t>& types, bool want_build_idx,
                              OutCol* outs, int32_t** pidx, int32_t** bidx, hipStream_t s, OutPlan& P) {
    const int nout = (int)types.size();
    const int64_t cap = out->capacity;
    size_t need = 0;
    auto place = [&](size_t bytes) {
        need = align_up(need, 256);
        size_t o = need;
        need += bytes;
        return o;
    };
    P.host = out->mem == VR_MEM_HOST;
    P.voff.assign(nout, 0);
    P.doff.assign(nout, 0);
    if (P.host) {
        if (out->probe_idx) P.pidx_off = place((size_t)cap * 4);
        if (out->build_idx && want_build_idx) P.bidx_off = place((size_t)cap * 4);
    }
    for (int q = 0; q < nout; q++) {
        const vr_out_column& oc = out->columns[q];
        if (oc.type != types[q]) return fail(VR_ERR_INVALID_ARGUMENT, "output column lane differs from its source");
        if (P.host) {
            if (oc.validity) P.voff[q] = place(align_up(bitmap_bytes(cap), 4));
            P.doff[q] = place(lane_bytes(types[q], cap));
        }
    }
    if (need) {
        VR_TRY(P.stage.ensure(j.dev, need));
        P.ob = P.stage.as<uint8_t>();
    }
    *pidx = P.host ? (out->probe_idx ? (int32_t*)(P.ob + P.pidx_off) : nullptr) : out->probe_idx;
    *bidx = want_build_idx ? (P.host ? (out->build_idx ? (int32_t*)(P.ob + P.bidx_off) : nullptr) : out->build_idx) : nullptr;
    for (int q = 0; q < nout; q++) {
        const vr_out_column& oc = out->columns[q];
        OutCol& d = outs[q];
        d.type = oc.type;
        if (P.host) {
            d.valid = oc.validity ? P.ob + P.voff[q] : nullptr;
            d.data = P.ob + P.doff[q];
        } else {
            d.valid = oc.validity;
            d.data = oc.data;
        }
        if (d.valid) {
            uint64_t words = (bitmap_bytes(cap) + 7) / 8;  // output bitmaps are padded to 8 bytes
            hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(words, 256, 1024)), dim3(256), 0, s, (uint64_t*)d.valid, words);
        }
    }
    return VR_OK;
}

// Copies `written` rows of staged outputs back to the caller's host buffers.
static vr_status copy_back(const vr_probe_output* out, const OutPlan& P, int64_t written, bool build_idx) {
    if (!P.host || written <= 0) return VR_OK;
    if (out->probe_idx) VR_TRY_HIP(hipMemcpy(out->probe_idx, P.ob + P.pidx_off, (size_t)written * 4, hipMemcpyDeviceToHost));
    if (out->build_idx && build_idx)
        VR_TRY_HIP(hipMemcpy(out->build_idx, P.ob + P.bidx_off, (size_t)written * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < P.doff.size(); q++) {
        const vr_out_column& oc = out->columns[q];
        if (oc.validity) VR_TRY_HIP(hipMemcpy(oc.validity, P.ob + P.voff[q], bitmap_bytes(written), hipMemcpyDeviceToHost));
        VR_TRY_HIP(hipMemcpy(oc.data, P.ob + P.doff[q], lane_bytes(oc.type, written), hipMemcpyDeviceToHost));
    }
    return VR_OK;
}

// What a probe call reads and writes, checked against the table and the output description: the
// streamed column (index into a batch's `columns`) of each slot, and the output lanes.
struct ProbeShape {
    std::vector<int32_t> slot_col;
    std::vector<int32_t> out_types;
    bool pairs = false;
};

// The checks of a probe call that need no streamed columns.
static vr_status check_probe_call(const JoinImpl& j, int32_t join_type, const vr_probe_output* out) {
    if (!j.finished) return fail(VR_ERR_STATE, "probe before vr_join_build_finish");
    if (join_type < VR_JOIN_INNER || join_type > VR_JOIN_OUTER) return fail(VR_ERR_INVALID_ARGUMENT, "join type");
    const bool pairs = join_type == VR_JOIN_INNER || join_type == VR_JOIN_OUTER;
    if (!pairs && out->num_build_out) return fail(VR_ERR_INVALID_ARGUMENT, "semi/anti joins output streamed columns only");
    if (out->num_probe_out < 0 || out->num_build_out < 0 || out->num_probe_out + out->num_build_out > 16 ||
        out->num_probe_out > 8 || out->num_build_out > 8)
        return fail(VR_ERR_UNSUPPORTED, "at most 8 + 8 output columns");
    if (join_type == VR_JOIN_OUTER)
        for (int q = 0; q < out->num_build_out; q++)
            if (!out->columns[out->num_probe_out + q].validity)
                return fail(VR_ERR_INVALID_ARGUMENT, "an outer join's build columns need validity for the padded rows");
    return VR_OK;
}

// Fills everything of the launch arguments but the batches (after check_probe_call).
static vr_status prepare_probe(JoinImpl& j, const vr_column* columns, int32_t num_columns, const int32_t* key_columns,
                               int32_t join_type, int32_t drop_null_keys, const vr_probe_output* out, ProbeArgs& a,
                               ProbeShape& sh) {
    const bool pairs = join_type == VR_JOIN_INNER || join_type == VR_JOIN_OUTER;
    sh.pairs = pairs;
    a.t = JoinTable{j.buckets.as<uint64_t>(), j.nbuckets, j.range_rows.as<uint32_t>(), j.sentinel_value, j.wide ? 1 : 0, 0};
    if (j.wide) a.bk = wide_keys(j);
    a.dense = DenseIndex{j.dense_range ? j.dense.as<uint64_t>() : nullptr, j.dense_min, j.dense_range};
    a.matched = (pairs && j.track) ? j.matched.as<uint32_t>() : nullptr;
    a.nk = (int)j.key_types.size();
    a.join_type = join_type;
    a.drop_null_keys = drop_null_keys;
    a.build_rows = j.rows > 0 ? 1 : 0;
    a.capacity = out->capacity;
    auto add_slot = [&](int32_t col, int32_t type) {
        a.ctype[sh.slot_col.size()] = type;
        sh.slot_col.push_back(col);
        return (int32_t)sh.slot_col.size() - 1;
    };
    for (int q = 0; q < a.nk; q++) {
        const int kc = key_columns[q];
        if (kc < 0 || kc >= num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "probe key column out of range");
        if (columns[kc].type != j.key_types[q]) return fail(VR_ERR_INVALID_ARGUMENT, "probe key lane differs from the build key");
        add_slot(kc, columns[kc].type);
    }
    a.nprobe_out = out->num_probe_out;
    a.nbuild_out = out->num_build_out;
    for (int q = 0; q < a.nprobe_out; q++) {
        const int pc = out->probe_out_columns[q];
        if (pc < 0 || pc >= num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "probe output column out of range");
        const int32_t t = columns[pc].type;
        if (t == VR_UTF8 || t == VR_BOOL) return fail(VR_ERR_UNSUPPORTED, "UTF8/BOOL outputs run on the CPU kernels");
        add_slot(pc, t);
        sh.out_types.push_back(t);
    }
    for (int q = 0; q < a.nbuild_out; q++) {
        const int bc = out->build_out_columns[q];
        if (bc < 0 || bc >= (int)j.cols.size()) return fail(VR_ERR_INVALID_ARGUMENT, "build output column out of range");
        if (j.col_types[bc] == VR_UTF8) return fail(VR_ERR_UNSUPPORTED, "UTF8/BOOL outputs run on the CPU kernels");
        a.build_src[q].type = j.col_types[bc];
        a.build_src[q].data = j.cols[bc].data.p;
        a.build_src[q].valid = j.cols[bc].has_valid ? j.cols[bc].valid.as<uint8_t>() : nullptr;
        a.build_from_key[q] = -1;
        for (int k = 0; k < a.nk; k++)
            if (j.key_cols[k] == bc) a.build_from_key[q] = k;  // a matched pair's build key = its streamed key
        sh.out_types.push_back(j.col_types[bc]);
    }
    if (const vr_join_residual* res = out->residual) {
        if (res->op < VR_CMP_EQ || res->op > VR_CMP_GE) return fail(VR_ERR_INVALID_ARGUMENT, "residual operator");
        if (drop_null_keys) return fail(VR_ERR_UNSUPPORTED, "a null-aware anti join with a residual runs on the CPU kernels");
        int32_t types[2];
        DCol* dst[2] = {&a.res_lbuild, &a.res_rbuild};
        int32_t* slot[2] = {&a.res_lslot, &a.res_rslot};
        const int32_t side[2] = {res->left_side, res->right_side}, col[2] = {res->left_column, res->right_column};
        for (int q = 0; q < 2; q++) {
            if (side[q] == 0) {
                if (col[q] < 0 || col[q] >= num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "residual column out of range");
                types[q] = columns[col[q]].type;
                *slot[q] = add_slot(col[q], types[q]);
            } else if (side[q] == 1) {
                if (col[q] < 0 || col[q] >= (int)j.cols.size()) return fail(VR_ERR_INVALID_ARGUMENT, "residual column out of range");
                types[q] = j.col_types[col[q]];
                dst[q]->type = types[q];
                dst[q]->data = j.cols[col[q]].data.p;
                dst[q]->valid = j.cols[col[q]].has_valid ? j.cols[col[q]].valid.as<uint8_t>() : nullptr;
            } else {
                return fail(VR_ERR_INVALID_ARGUMENT, "residual side is 0 (streamed) or 1 (build)");
            }
        }
        if (types[0] != types[1] || (types[0] != VR_INT32 && types[0] != VR_INT64))
            return fail(VR_ERR_UNSUPPORTED, "residuals over other than two INT32 or two INT64 lanes run on the CPU kernels");
        a.res_on = 1;
        a.res_op = res->op;
        a.res_lb = side[0];
        a.res_rb = side[1];
        a.res_loff = res->left_offset;
        a.res_roff = res->right_offset;
    }
    // the payload table, when every build column the probe writes is a key or in its cells
    for (int q = 0; q < 8; q++) a.pay_off[q] = -1;
    bool use_pay = j.pay_nb && pairs && !a.res_on, needs_pay = false;
    for (int q = 0; q < a.nbuild_out && use_pay; q++) {
        const int bc = out->build_out_columns[q];
        if (a.build_from_key[q] < 0 && j.pay_off[bc] < 0) use_pay = false;
        a.pay_off[q] = a.build_from_key[q] < 0 ? j.pay_off[bc] : -1;
        if (a.pay_off[q] >= 0) needs_pay = true;
    }
    use_pay = use_pay && needs_pay;  // build keys alone come from the streamed keys
    if (use_pay) {
        a.pt = PayTable{j.pay.as<uint64_t>(), j.pay_nb};
        a.pay_words = j.pay_words;
        j.dev->count(&vr_stats::payload_join_probes, 1);
    } else {
        a.pt = PayTable{nullptr, 0};
        for (int q = 0; q < 8; q++) a.pay_off[q] = -1;
    }
    a.ncb = (int32_t)sh.slot_col.size();
    return VR_OK;
}

vr_status vr_join_probe(vr_join* h, const vr_column* columns, int32_t num_columns, const int32_t* key_columns, int64_t n,
                        const uint8_t* selection, int32_t join_type, int32_t drop_null_keys, vr_probe_output* out,
                        void* stream) {
    if (!h || !out || (!columns && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    VR_TRY(check_probe_call(j, join_type, out));
    out->rows = 0;
    if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative row count");
    if (n == 0) return VR_OK;
    if (n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "probe batches past 2^31 rows");
    ProbeArgs a{};
    ProbeShape sh;
    VR_TRY(prepare_probe(j, columns, num_columns, key_columns, join_type, drop_null_keys, out, a, sh));
    hipStream_t s = j.dev->pick(stream);
    a.one = PBatch{n, 0, selection, 0};
    // streamed columns: device pointers, or staged from host memory
    std::vector<StagePiece> pieces;
    auto want = [&](const void* p, size_t bytes, const void** dst, int32_t mem) {
        *dst = p;
        if (p && bytes && mem == VR_MEM_HOST) pieces.push_back({p, bytes, dst});
    };
    std::vector<DevBuf> key_ids;  // UTF8 keys: the rows' ids, read by the probe kernel
    for (size_t q = 0; q < sh.slot_col.size(); q++) {
        const vr_column& c = columns[sh.slot_col[q]];
        want(c.validity, bitmap_bytes(n), (const void**)&a.one_cols[q].valid, c.mem);
        if (q < (size_t)a.nk && c.type == VR_UTF8) {
            key_ids.emplace_back();
            VR_TRY(key_ids.back().ensure(j.dev, (size_t)n * 4));
            VR_TRY(map_key_strings(j, (int)q, c, n, false, key_ids.back().as<int32_t>(), s));
            a.one_cols[q].data = key_ids.back().p;
        } else {
            want(c.data, lane_bytes(c.type, n), &a.one_cols[q].data, c.mem);
        }
    }
    if (selection && columns[key_columns[0]].mem == VR_MEM_HOST)
        want(selection, bitmap_bytes(n), (const void**)&a.one.sel, VR_MEM_HOST);
    std::unique_lock<std::mutex> stage_lock(j.stage_mu, std::defer_lock);
    int slot = -1;
    if (!pieces.empty()) {
        stage_lock.lock();  // the shared pinned staging, until the kernel reading it has finished
        VR_TRY(j.staging.stage(j.dev, pieces, s, &slot));
    }
    OutPlan P;
    VR_TRY(plan_outputs(j, out, sh.out_types, sh.pairs, a.outs, &a.probe_idx, &a.build_idx, s, P));
    const int64_t cap = out->capacity;
    // tiles and their look-back state, the call's own
    a.ntiles = (n + PROBE_TILE - 1) / PROBE_TILE;
    DevBuf tiles;
    VR_TRY(tiles.ensure(j.dev, (size_t)(a.ntiles + 2) * 8));
    a.tile_counter = tiles.as<uint64_t>();
    a.tile_state = tiles.as<uint64_t>() + 2;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(a.ntiles + 2, 256, 1024)), dim3(256), 0, s, tiles.as<uint64_t>(),
                       (uint64_t)(a.ntiles + 2));
    if (a.pt.buckets) hipLaunchKernelGGL((join_probe_kernel<false, true>), dim3((unsigned)a.ntiles), dim3(PROBE_BLOCK), pay_lds(a.pay_words), s, a);
    else hipLaunchKernelGGL((join_probe_kernel<false, false>), dim3((unsigned)a.ntiles), dim3(PROBE_BLOCK), 0, s, a);
    VR_TRY_HIP(hipGetLastError());
    uint64_t rows = 0;
    VR_TRY_HIP(hipMemcpyAsync(&rows, tiles.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (slot >= 0) VR_TRY(j.staging.done(slot, s));
    if (stage_lock.owns_lock()) stage_lock.unlock();
    int64_t written = std::min<int64_t>((int64_t)rows, cap);
    VR_TRY(copy_back(out, P, written, sh.pairs));
    out->rows = (int64_t)rows;
    j.dev->count(&vr_stats::calls, 1);
    j.dev->count(&vr_stats::input_rows, n);
    j.dev->count(&vr_stats::output_rows, written);
    j.dev->count(&vr_stats::kernel_launches, 2);
    if ((int64_t)rows > cap) return fail(VR_ERR_CAPACITY, "probe output needs " + std::to_string(rows) + " rows");
    return VR_OK;
}

vr_status vr_join_probe_noalloc(vr_join* h, const vr_column* columns, int32_t num_columns, const int32_t* key_columns,
                                int64_t n, const uint8_t* selection, int32_t join_type, int32_t drop_null_keys,
                                vr_probe_output* out, void* stream) {
    if (!h || !out || (!columns && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    VR_TRY(check_probe_call(j, join_type, out));
    out->rows = 0;
    if (out->mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "the allocation-free probe writes device outputs only");
    if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative row count");
    if (n == 0) return VR_OK;
    const int64_t ntiles = (n + PROBE_TILE - 1) / PROBE_TILE;
    if ((uint64_t)ntiles + 2 > NOALLOC_SCRATCH_WORDS)
        return fail(VR_ERR_UNSUPPORTED, "the allocation-free probe takes at most " +
                                            std::to_string((NOALLOC_SCRATCH_WORDS - 2) * PROBE_TILE) + " rows per call");
    ProbeArgs a{};
    ProbeShape sh;
    VR_TRY(prepare_probe(j, columns, num_columns, key_columns, join_type, drop_null_keys, out, a, sh));
    for (size_t q = 0; q < sh.slot_col.size(); q++) {
        const vr_column& c = columns[sh.slot_col[q]];
        if (c.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "the allocation-free probe reads device-resident columns only");
        if (c.type == VR_UTF8) return fail(VR_ERR_UNSUPPORTED, "the allocation-free probe takes no UTF8 keys (their ids need memory)");
        a.one_cols[q] = PCol{c.validity, c.data};
    }
    a.one = PBatch{n, 0, selection, 0};
    VR_TRY_HIP(hipSetDevice(j.dev->ordinal));
    uint64_t* sp = nullptr;
    VR_TRY(scratch_on(j.dev->ordinal, &sp));
    hipStream_t s = j.dev->pick(stream);
    OutPlan P;  // device outputs: no staging
    VR_TRY(plan_outputs(j, out, sh.out_types, sh.pairs, a.outs, &a.probe_idx, &a.build_idx, s, P));
    a.ntiles = ntiles;
    a.tile_counter = sp;
    a.tile_state = sp + 2;
    // cleared by a kernel, not a memset: the call may be captured into a graph
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(ntiles + 2, 256, 1024)), dim3(256), 0, s, sp, (uint64_t)(ntiles + 2));
    if (a.pt.buckets) hipLaunchKernelGGL((join_probe_kernel<false, true>), dim3((unsigned)ntiles), dim3(PROBE_BLOCK), pay_lds(a.pay_words), s, a);
    else hipLaunchKernelGGL((join_probe_kernel<false, false>), dim3((unsigned)ntiles), dim3(PROBE_BLOCK), 0, s, a);
    VR_TRY_HIP(hipGetLastError());
    uint64_t rows = 0;
    VR_TRY_HIP(hipMemcpyAsync(&rows, sp + 1, 8, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    out->rows = (int64_t)rows;
    j.dev->count(&vr_stats::calls, 1);
    j.dev->count(&vr_stats::input_rows, n);
    j.dev->count(&vr_stats::output_rows, std::min<int64_t>((int64_t)rows, out->capacity));
    j.dev->count(&vr_stats::kernel_launches, 2);
    if ((int64_t)rows > out->capacity) return fail(VR_ERR_CAPACITY, "probe output needs " + std::to_string(rows) + " rows");
    return VR_OK;
}

vr_status vr_join_probe_batches(vr_join* h, int32_t num_batches, const vr_column* columns, int32_t num_columns,
                                const int64_t* lengths, const uint8_t* const* selections, const int32_t* key_columns,
                                int32_t join_type, int32_t drop_null_keys, vr_probe_output* out, int64_t* batch_rows,
                                void* stream) {
    if (!h || !out || num_batches < 0 || (num_batches && (!columns || !lengths))) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    VR_TRY(check_probe_call(j, join_type, out));
    out->rows = 0;
    if (num_batches == 0) return VR_OK;
    if (out->mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "batched probes write device outputs (vr_join_probe stages host ones)");
    ProbeArgs a{};
    ProbeShape sh;
    VR_TRY(prepare_probe(j, columns, num_columns, key_columns, join_type, drop_null_keys, out, a, sh));
    const int ncb = a.ncb;
    std::vector<PBatch> batches((size_t)num_batches);
    std::vector<PCol> bcols((size_t)num_batches * ncb);
    int64_t tiles = 0, rows_in = 0;
    hipStream_t s = j.dev->pick(stream);
    std::vector<DevBuf> key_ids;  // UTF8 keys: each batch's ids, read by the probe kernel
    for (int32_t b = 0; b < num_batches; b++) {
        const int64_t n = lengths[b];
        if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative row count");
        if (n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "probe batches past 2^31 rows");
        const vr_column* cb = columns + (size_t)b * num_columns;
        for (int q = 0; q < ncb; q++) {
            const vr_column& c = cb[sh.slot_col[q]];
            if (c.type != a.ctype[q]) return fail(VR_ERR_INVALID_ARGUMENT, "a batch's column lane differs from the first batch's");
            if (n && c.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "batched probes read device-resident batches");
            if (n && c.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "a batch's column is shorter than the batch");
            bcols[(size_t)b * ncb + q] = PCol{c.validity, c.data};
            if (n && q < a.nk && c.type == VR_UTF8) {
                key_ids.emplace_back();
                VR_TRY(key_ids.back().ensure(j.dev, (size_t)n * 4));
                VR_TRY(map_key_strings(j, q, c, n, false, key_ids.back().as<int32_t>(), s));
                bcols[(size_t)b * ncb + q].data = key_ids.back().p;
            }
        }
        batches[b] = PBatch{n, tiles, selections ? selections[b] : nullptr, 0};
        tiles += (n + PROBE_TILE - 1) / PROBE_TILE;
        rows_in += n;
    }
    if (tiles == 0) {
        if (batch_rows) std::fill(batch_rows, batch_rows + num_batches, 0);
        return VR_OK;
    }
    if (tiles >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "more than 2^31 probe tiles in one call");
    // the call's device tables: batches, their columns, tile -> batch, tile states, batch ends
    const size_t off_cols = align_up(sizeof(PBatch) * (size_t)num_batches, 256);
    const size_t off_map = off_cols + align_up(sizeof(PCol) * bcols.size(), 256);
    const size_t off_state = off_map + align_up(4 * (size_t)tiles, 256);
    const size_t off_end = off_state + align_up(8 * (size_t)(tiles + 2), 256);
    const size_t bytes = off_end + 8 * (size_t)num_batches;
    DevBuf desc;
    VR_TRY(desc.ensure(j.dev, bytes));
    uint8_t* d = desc.as<uint8_t>();
    VR_TRY_HIP(hipMemcpyAsync(d, batches.data(), sizeof(PBatch) * batches.size(), hipMemcpyHostToDevice, s));
    VR_TRY_HIP(hipMemcpyAsync(d + off_cols, bcols.data(), sizeof(PCol) * bcols.size(), hipMemcpyHostToDevice, s));
    a.batches = (const PBatch*)d;
    a.bcols = (const PCol*)(d + off_cols);
    a.tile_batch = (const int32_t*)(d + off_map);
    a.tile_counter = (uint64_t*)(d + off_state);
    a.tile_state = a.tile_counter + 2;
    a.batch_end = (int64_t*)(d + off_end);
    a.ntiles = tiles;
    OutPlan P;
    VR_TRY(plan_outputs(j, out, sh.out_types, sh.pairs, a.outs, &a.probe_idx, &a.build_idx, s, P));
    const int cap = j.dev->num_cus * 16;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_of(tiles + 2, 256, cap)), dim3(256), 0, s, a.tile_counter, (uint64_t)(tiles + 2));
    hipLaunchKernelGGL(tile_batch_kernel, dim3(grid_of(num_batches, 256, cap)), dim3(256), 0, s, a.batches, num_batches,
                       (int32_t*)a.tile_batch);
    if (a.pt.buckets) hipLaunchKernelGGL((join_probe_kernel<true, true>), dim3((unsigned)tiles), dim3(PROBE_BLOCK), pay_lds(a.pay_words), s, a);
    else hipLaunchKernelGGL((join_probe_kernel<true, false>), dim3((unsigned)tiles), dim3(PROBE_BLOCK), 0, s, a);
    VR_TRY_HIP(hipGetLastError());
    uint64_t rows = 0;
    std::vector<int64_t> ends(batch_rows ? (size_t)num_batches : 0);
    VR_TRY_HIP(hipMemcpyAsync(&rows, a.tile_counter + 1, 8, hipMemcpyDeviceToHost, s));
    if (batch_rows) VR_TRY_HIP(hipMemcpyAsync(ends.data(), a.batch_end, 8 * (size_t)num_batches, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (batch_rows) {  // an empty batch wrote no end: it ends where the batch before it ends
        int64_t prev = 0;
        for (int32_t b = 0; b < num_batches; b++) {
            const int64_t e = batches[b].n ? ends[b] : prev;
            batch_rows[b] = e - prev;
            prev = e;
        }
    }
    out->rows = (int64_t)rows;
    const int64_t written = std::min<int64_t>((int64_t)rows, out->capacity);
    j.dev->count(&vr_stats::calls, 1);
    j.dev->count(&vr_stats::input_rows, rows_in);
    j.dev->count(&vr_stats::output_rows, written);
    j.dev->count(&vr_stats::kernel_launches, 3);
    if ((int64_t)rows > out->capacity) return fail(VR_ERR_CAPACITY, "probe output needs " + std::to_string(rows) + " rows");
    return VR_OK;
}

vr_status vr_join_unmatched_build(vr_join* h, vr_probe_output* out, void* stream) {
    if (!h || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    if (!j.finished) return fail(VR_ERR_STATE, "unmatched rows before vr_join_build_finish");
    if (!j.track) return fail(VR_ERR_STATE, "the join does not track matched build rows (track_matched)");
    if (out->num_probe_out + out->num_build_out > 16 || out->num_probe_out > 8 || out->num_build_out > 8)
        return fail(VR_ERR_UNSUPPORTED, "at most 8 + 8 output columns");
    for (int q = 0; q < out->num_probe_out; q++)
        if (!out->columns[q].validity) return fail(VR_ERR_INVALID_ARGUMENT, "the null streamed columns need validity");
    out->rows = 0;
    hipStream_t s = j.dev->pick(stream);
    const int64_t n = j.rows;
    // the unmatched rows in build order
    VR_TRY(j.unmatched.ensure(j.dev, (size_t)std::max<int64_t>(n, 1) * 4 + 8));
    int32_t* list = j.unmatched.as<int32_t>();
    int64_t* count = (int64_t*)(j.unmatched.as<uint8_t>() + (size_t)std::max<int64_t>(n, 1) * 4);
    int64_t found = 0;
    if (n > 0) {
        size_t tmp = 0;
        hipcub::CountingInputIterator<int32_t> rows_in(0);
        const UnmatchedRow pick{j.matched.as<uint32_t>()};
        VR_TRY_HIP(hipcub::DeviceSelect::If(nullptr, tmp, rows_in, list, count, (int)n, pick, s));
        VR_TRY(j.cub_tmp.ensure(j.dev, tmp));
        VR_TRY_HIP(hipcub::DeviceSelect::If(j.cub_tmp.p, tmp, rows_in, list, count, (int)n, pick, s));
        VR_TRY_HIP(hipMemcpyAsync(&found, count, 8, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipStreamSynchronize(s));
    }
    out->rows = found;
    if (found > out->capacity) return fail(VR_ERR_CAPACITY, "unmatched build rows need " + std::to_string(found) + " rows");
    if (found == 0) return VR_OK;
    UnmatchedArgs a{};
    a.rows = list;
    a.m = found;
    a.nprobe_out = out->num_probe_out;
    a.nbuild_out = out->num_build_out;
    std::vector<int32_t> types;
    for (int q = 0; q < a.nprobe_out; q++) types.push_back(out->columns[q].type);
    for (int q = 0; q < a.nbuild_out; q++) {
        int bc = out->build_out_columns[q];
        if (bc < 0 || bc >= (int)j.cols.size()) return fail(VR_ERR_INVALID_ARGUMENT, "build output column out of range");
        if (j.col_types[bc] == VR_UTF8) return fail(VR_ERR_UNSUPPORTED, "UTF8/BOOL outputs run on the CPU kernels");
        a.build_src[q].type = j.col_types[bc];
        a.build_src[q].data = j.cols[bc].data.p;
        a.build_src[q].valid = j.cols[bc].has_valid ? j.cols[bc].valid.as<uint8_t>() : nullptr;
        types.push_back(j.col_types[bc]);
    }
    OutPlan P;
    VR_TRY(plan_outputs(j, out, types, true, a.outs, &a.probe_idx, &a.build_idx, s, P));
    hipLaunchKernelGGL(unmatched_rows_kernel, dim3(grid_of((uint64_t)found, 256, j.dev->num_cus * 16)), dim3(256), 0, s, a);
    VR_TRY_HIP(hipGetLastError());
    VR_TRY_HIP(hipStreamSynchronize(s));
    VR_TRY(copy_back(out, P, found, true));
    j.dev->count(&vr_stats::calls, 1);
    j.dev->count(&vr_stats::output_rows, found);
    j.dev->count(&vr_stats::kernel_launches, 3);
    return VR_OK;
}

vr_status vr_join_memory(vr_join* h, int64_t* bytes) {
    if (!h || !bytes) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    JoinImpl& j = h->j;
    int64_t b = (int64_t)(j.buckets.bytes + j.range_rows.bytes + j.cub_tmp.bytes + j.small.bytes + j.dense.bytes + j.pay.bytes +
                          j.matched.bytes + j.unmatched.bytes);
    for (auto& c : j.cols) b += (int64_t)(c.data.bytes + c.valid.bytes);
    for (auto& t : j.strs) b += t.memory();
    *bytes = b;
    return VR_OK;
}

vr_status vr_join_destroy(vr_join* h) {
    if (!h) return VR_OK;
    (void)hipStreamSynchronize(h->j.dev->stream);
    delete h;
    return VR_OK;
}

}  // extern "C"
