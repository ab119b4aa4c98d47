// Partitioned group-by update for keys of at most 64 bits: the batched path without per-row
// device atomics.
//
// Device-scope atomics on this part run at ~24 G/s whatever the working set, while LDS atomics run
// at ~2 T/s and a wave streams a block of 64 records at HBM speed.  The update is therefore split
// into launches in which no global state row is ever updated by two workgroups:
//
//   A1 (count)   Each workgroup walks the tiles its scatter workgroup will walk, reading only the
//                keys (and the selection and filter), and counts per region the rows that will
//                become records: rows whose key is not in the hot set (the most frequent keys of a
//                sample of the input).  Two small scans turn the counts into a range of the
//                record pool per (workgroup, region), the ranges of a region contiguous and each
//                region starting on a block of PART_BLK records.
//   A2 (scatter) Each workgroup streams its tiles once.  Hot rows fold into a per-workgroup LDS
//                copy of the hot groups; every other row becomes a record -- the key, the value
//                lanes the aggregates read, the row index when a double MIN/MAX needs it, one
//                validity word -- written at its region's LDS cursor inside the workgroup's
//                range.  The waves never wait for each other inside the tile loop.
//   B (merge)    One workgroup per region owns that region's slots and state rows.  It first
//                merges the hot set's partials (reduced over the scatter workgroups) of its region,
//                then its waves take the region's record blocks from an LDS counter (no barrier),
//                folding records into an LDS cache of groups with LDS atomics, and at the end the
//                cache is merged into the region -- new slots claimed without device atomics, state
//                rows by plain read-modify-write, since no other workgroup touches the region.  A
//                record whose key finds the cache full around it is marked in a pending map and
//                folded by a further pass after that merge.
//
// Growth.  Before a merge adds groups to its region the workgroup checks that the region can
// take every cached group as a new one; when it cannot, the workgroup stops and leaves in
// resume_step the pass to redo (0: the hot partials too).  The pending maps are used in turn, so
// the stopped pass's input is intact; the table grows (regions double, their number is kept) and
// the stopped regions resume there.
#pragma once

#include "vr_agg_kernels.h"
#include "vr_part_types.h"

namespace vr {

// Stand-ins read when a batch has no bitmap, selection or filter, so that every load of a tile is
// issued whatever the batch holds (no branch between a load and the next).
__device__ const uint8_t k_all_ones[8] = {0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF};
__device__ __align__(16) const uint32_t k_zero_words[4] = {0, 0, 0, 0};

struct PartArgs {
    UpdateArgs u;
    RecLayout rl;
    ColProg prog[VR_MAX_VALUES];
    int32_t nprog;
    int32_t count_split;    // count-pass workgroups per scatter workgroup (each a share of its tiles)
    int32_t nparts;         // regions = 2^u.t.pbits
    int32_t nhot;           // hot-set entries (unused entries hold EMPTY_KEY)
    int32_t hot_slots;      // LDS index slots of the hot set (power of two >= 2 * nhot)
    int32_t grid_a;         // scatter workgroups = chains per region
    int32_t cache_e;        // merge: LDS cache entries (power of two, >= 2 * PB_BLOCK)
    int32_t resumed;        // merge: start every region at resume_step
    uint32_t region_limit;  // groups a region may hold
    int32_t pad;
    const uint64_t* hot_keys;  // [nhot]
    uint64_t* hot_part;        // [grid_a][nhot][stride]
    uint64_t* hot_total;       // [nhot][stride]
    uint32_t* pool;            // block b holds records [b * 64, b * 64 + 64) of rl.words words
    uint32_t* count;           // [count_split][grid_a][nparts] records of each scatter workgroup per
                               // region, by count workgroup (the column scan leaves the totals in
                               // the first [grid_a][nparts])
    uint32_t* base;            // [grid_a][nparts] where they start inside the region's records
    uint32_t* rtot;            // [nparts] records per region
    uint32_t* rstart;          // [nparts + 1] first record of each region (a multiple of PART_BLK)
    int32_t* resume_step;      // [nparts] merge: see above; -1 when merged
    const uint32_t* tile_batch;  // scatter: the batch of every tile
    uint64_t* pending;         // merge: 2 x pool_blocks words, the records of a block pending in a pass
    uint64_t pool_blocks;      // blocks of the pool
    uint32_t pool_records;     // records the pool holds (pool_blocks * PART_BLK)
    uint32_t pad3;
};

// Where the kernels read the column programs and the record layout.  RuntimeProgs reads them from
// the launch arguments (one build serves every spec); a policy with the same members whose values
// are compile-time constants -- generated per spec and compiled at run time -- lets the compiler
// fold the programs into the per-row code (no per-row flag tests or register selects).
struct RuntimeProgs {
    static constexpr int max_rpt = PA_MAX_RPT;
    static __device__ __forceinline__ int nprog(const PartArgs& a) { return a.nprog; }
    static __device__ __forceinline__ const ColProg& prog(const PartArgs& a, int q) { return a.prog[q]; }
    static __device__ __forceinline__ int valid_off(const PartArgs& a) { return a.rl.valid_off; }
    static __device__ __forceinline__ int row_off(const PartArgs& a) { return a.rl.row_off; }
    // the key is one 64-bit lane (loaded with the tile; other key modes load theirs per row)
    static __device__ __forceinline__ bool fast_key(const PartArgs& a) { return a.u.ks.mode == KM_I64; }
    // rows per thread per tile (at most max_rpt)
    static __device__ __forceinline__ int rpt(const PartArgs& a) { return a.u.rows_per_thread; }
};

// Folds one value of a column into the state words at `e` (an LDS entry: LDS atomics), as its
// program says.  v: valid; x0/x1: the lane's words (INT32 in the low half of x0).
__device__ __forceinline__ void apply_prog(const UpdateArgs& u, const ColProg& cp, bool v, uint64_t x0, uint64_t x1,
                                           uint64_t* e) {
    // Lanes differ only in v and the values: every atomic below is issued by every lane, a null
    // row with the operation's neutral value, so that the wave never splits (the flags are uniform).
    const uint16_t fl = cp.flags;
    if (fl & CP_NULLS) atomic_add_u64(e + cp.nulls_off, v ? 0ull : 1ull);
    if (fl & CP_SUM_I64) {
        int64_t x = cp.type == VR_INT32 ? (int64_t)(int32_t)(uint32_t)x0 : (int64_t)x0;
        x = v ? x : 0;
        atomic_add_u64(e + cp.sum_off, (uint64_t)(uint32_t)x);
        atomic_add_u64(e + cp.sum_off + 1, (uint64_t)(x >> 32));
    }
    if (fl & CP_SUM_D128) {
        const uint64_t a0 = v ? x0 : 0, a1 = v ? x1 : 0;
        atomic_add_u64(e + cp.sum_off, a0 & 0xFFFFFFFFull);
        atomic_add_u64(e + cp.sum_off + 1, a0 >> 32);
        atomic_add_u64(e + cp.sum_off + 2, a1 & 0xFFFFFFFFull);
        atomic_add_u64(e + cp.sum_off + 3, (uint64_t)((int64_t)a1 >> 32));
    }
    if (fl & CP_SUM_F64) atomic_add_f64(e + cp.sum_off, v ? __longlong_as_double((long long)x0) : -0.0);
    if (fl & (CP_MIN_I32 | CP_MAX_I32)) {
        const int64_t x = (int64_t)(int32_t)(uint32_t)x0;
        if (fl & CP_MIN_I32) atomic_min_i64(e + cp.min_off, v ? x : INT64_MAX);
        if (fl & CP_MAX_I32) atomic_max_i64(e + cp.max_off, v ? x : INT64_MIN);
    }
    if ((fl & (CP_MIN_I64 | CP_MAX_I64)) && v) {
        const uint64_t o = x0 ^ 0x8000000000000000ull;
        uint64_t* f = e + u.flags_off;
        if (fl & CP_MIN_I64) {
            const uint64_t bit = 1ull << cp.min_bit;
            const bool has = (*f & bit) != 0;
            if (!has || o < e[cp.min_off]) atomic_min_u64(e + cp.min_off, o);
            if (!has) atomicOr((unsigned long long*)f, bit);
        }
        if (fl & CP_MAX_I64) {
            const uint64_t bit = 1ull << cp.max_bit;
            const bool has = (*f & bit) != 0;
            if (!has || o > e[cp.max_off]) atomic_max_u64(e + cp.max_off, o);
            if (!has) atomicOr((unsigned long long*)f, bit);
        }
    }
}

// Word o of a record held in registers (o uniform): a select per word, no indexed register access.
template <int RW>
__device__ __forceinline__ uint32_t get_word(const uint32_t (&w)[RW], int o) {
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < RW; j++)
        if (j == o) x = w[j];
    return x;
}

__device__ __forceinline__ const BatchDesc& batch_of_tile(const UpdateArgs& u, int64_t tile) {
    if (u.num_batches <= 1) return u.one;
    int lo = 0, hi = u.num_batches - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (batch_at(u.batches, mid).tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return batch_at(u.batches, lo);
}

// ------------------------------------------------------------------ hot set

// Counts the keys of HS_ROWS consecutive rows at the start of `nsamp` evenly spaced tiles (tiles
// of PA_BLOCK * rows_per_thread rows): each workgroup counts its sample in LDS, then adds each
// distinct key's count to `cand` (open addressing over cand_mask + 1 entries of key, count).  Keys
// past a full LDS or candidate table are dropped: the hot set is only a hint.
__global__ __launch_bounds__(HS_BLOCK) void hot_sample_kernel(UpdateArgs u, int64_t nsamp, uint64_t* cand,
                                                              uint32_t* cand_count, uint64_t cand_mask) {
    __shared__ uint64_t lkeys[HS_SLOTS];
    __shared__ uint32_t lcount[HS_SLOTS];
    const int64_t T = (int64_t)u.rows_per_thread * PA_BLOCK;
    for (int64_t s = blockIdx.x; s < nsamp; s += gridDim.x) {
        for (int q = threadIdx.x; q < HS_SLOTS; q += blockDim.x) {
            lkeys[q] = EMPTY_KEY;
            lcount[q] = 0;
        }
        __syncthreads();
        const int64_t tile = u.tile_lo + s * ((u.total_tiles - u.tile_lo) / nsamp);
        const BatchDesc& b = batch_of_tile(u, tile);
        const int64_t row0 = (tile - b.tile_begin) * T;
        for (int r = threadIdx.x; r < HS_ROWS; r += blockDim.x) {
            const int64_t i = row0 + r;
            if (i >= b.n) break;
            if (b.sel && !bit_at(b.sel, i)) continue;
            uint64_t key, slot;
            if (!key64_of(u, b, i, &key, &slot) || key == EMPTY_KEY) continue;
            uint32_t pos = (uint32_t)fmix64(key) & (HS_SLOTS - 1);
            for (int q = 0; q < HS_SLOTS; q++) {
                const uint64_t old = atomicCAS((unsigned long long*)&lkeys[pos], (unsigned long long)EMPTY_KEY,
                                               (unsigned long long)key);
                if (old == EMPTY_KEY || old == key) {
                    atomicAdd(&lcount[pos], 1u);
                    break;
                }
                pos = (pos + 1) & (HS_SLOTS - 1);
            }
        }
        __syncthreads();
        for (int q = threadIdx.x; q < HS_SLOTS; q += blockDim.x) {
            const uint64_t key = lkeys[q];
            if (key == EMPTY_KEY) continue;
            uint64_t pos = fmix64(key) & cand_mask;
            for (uint64_t t = 0; t <= cand_mask; t++) {
                uint64_t c = cand[pos];
                if (c == EMPTY_KEY) c = atomic_cas_u64(&cand[pos], EMPTY_KEY, key);
                if (c == EMPTY_KEY || c == key) {
                    atomicAdd(&cand_count[pos], lcount[q]);
                    break;
                }
                pos = (pos + 1) & cand_mask;
            }
        }
        __syncthreads();
    }
}

// hist[c] = candidates seen c times (c capped at 255).
__global__ __launch_bounds__(256) void hot_hist_kernel(const uint32_t* cand_count, uint64_t n, uint32_t* hist) {
    __shared__ uint32_t lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c = cand_count[i];
        if (c) atomicAdd(&lh[c > 255 ? 255 : c], 1u);
    }
    __syncthreads();
    if (lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

// The hot set: the candidates seen at least `thr` >= 2 times, `thr` the smallest count whose
// candidates fit in max_hot.  hot_keys[] must hold EMPTY_KEY beforehand.
__global__ __launch_bounds__(256) void hot_pick_kernel(const uint64_t* cand, const uint32_t* cand_count, uint64_t n,
                                                       const uint32_t* hist, int max_hot, uint64_t* hot_keys,
                                                       uint32_t* nhot) {
    __shared__ uint32_t thr;
    if (threadIdx.x == 0) {
        uint32_t acc = 0, t = 256;
        for (int c = 255; c >= 2; c--) {
            if (acc + hist[c] > (uint32_t)max_hot) break;
            acc += hist[c];
            t = (uint32_t)c;
        }
        thr = t;
    }
    __syncthreads();
    const uint32_t t = thr;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t c = cand_count[i];
        if (c >= t) {
            uint32_t k = atomicAdd(nhot, 1u);
            if (k < (uint32_t)max_hot) hot_keys[k] = cand[i];
        }
    }
}

// ------------------------------------------------------------------ phase A: count and scatter

// LDS rows of the hot set and of the merge cache: the key, then the state words, padded to an odd
// number of 8-byte words so that the lanes of a wave updating one word of different rows spread
// over the LDS banks (a 64-byte row would put them all on two).
__host__ __device__ inline int lds_row_words(int stride) { return (stride + 1) | 1; }

// The hot set's LDS index: buckets of HB entries, an entry = 16 bits of the key's hash (bits
// 32-47, not the bucket's; never 0xFFFF, so no entry of a key reads as an empty one) above the
// key's number in the hot set.  Entries fill each bucket from its start, so a bucket with an empty
// entry ends the search; one bucket read answers almost every lookup.
constexpr int HB = 8;
__device__ __forceinline__ uint32_t hot_tag(uint64_t h) {
    const uint32_t t = (uint32_t)(h >> 16) & 0xFFFF0000u;
    return t < 0xFFFF0000u ? t : 0xFFFE0000u;
}

// The entry of `key` among bucket b's entries whose tag matches (a tag collision): a call of its
// own, so that none of its work is hoisted into the common path.
__device__ __noinline__ uint32_t hot_find_collided(const uint32_t* hidx, const uint64_t* hrow, int HW, uint32_t b,
                                                   uint32_t tag, uint64_t key) {
    for (int j = 0; j < HB; j++) {
        const uint32_t e = hidx[b * HB + j];
        if ((e ^ tag) < 0x10000u && hrow[(size_t)(e & 0xFFFFu) * HW] == key) return e & 0xFFFFu;
    }
    return HOT_EMPTY;
}

// Number of `key` in the hot set (keys hrow[k * HW]), or HOT_EMPTY.  A bucket's entries are
// compared without branches (three instructions an entry: the tag's half of the entry against the
// key's, the last match kept); the one candidate's key is read back, and only a second key of the
// bucket with the same 16-bit tag (a tag collision) takes the slow loop.
__device__ __forceinline__ uint32_t hot_find(const uint32_t* hidx, const uint64_t* hrow, int HW, uint32_t bmask,
                                             uint64_t h, uint64_t key) {
    const uint32_t tag = hot_tag(h);
    uint32_t b = (uint32_t)h & bmask;
    for (uint32_t probes = 0; probes <= bmask; probes++) {
        const uint4* q = (const uint4*)(hidx + b * HB);
        const uint4 x = q[0], y = q[1];
        const uint32_t e[HB] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
        uint32_t cand = HOT_EMPTY;
#pragma unroll
        for (int j = HB - 1; j >= 0; j--) cand = (e[j] ^ tag) < 0x10000u ? e[j] : cand;  // the first match
        const uint32_t k = cand != HOT_EMPTY ? (cand & 0xFFFFu) : 0u;
        const bool hit = cand != HOT_EMPTY && hrow[(size_t)k * HW] == key;
        if (hit) return k;
        if (cand != HOT_EMPTY) {  // a tag collision: every matching entry's key
            const uint32_t c = hot_find_collided(hidx, hrow, HW, b, tag, key);
            if (c != HOT_EMPTY) return c;
        }
        if (e[HB - 1] == HOT_EMPTY) return HOT_EMPTY;  // the bucket has room: the key is not in the set
        b = (b + 1) & bmask;
    }
    return HOT_EMPTY;
}

// Enters the keys hrow[k * HW], k < K, into the hot set's LDS index (EMPTY_KEY entries are not in
// the set).  Called by every thread of the workgroup, between two barriers.
__device__ __forceinline__ void build_hot_index(uint32_t* hidx, const uint64_t* hrow, int HW, int K, uint32_t hbmask) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const uint64_t key = hrow[(size_t)k * HW];
        if (key == EMPTY_KEY) continue;
        const uint64_t h = fmix64(key);
        const uint32_t x = hot_tag(h) | (uint32_t)k;
        for (uint32_t b = (uint32_t)h & hbmask;; b = (b + 1) & hbmask) {  // K < H: a bucket has room
            int j = 0;
            while (j < HB && atomicCAS(&hidx[b * HB + j], HOT_EMPTY, x) != HOT_EMPTY) j++;
            if (j < HB) break;
        }
    }
}

// The descriptor of a tile's batch, through the constant address space: the tile number is the
// same in every lane, so the batch number and the descriptor's fields come in scalar loads.
__device__ __forceinline__ const BatchDesc& tile_desc(const PartArgs& a, int64_t tile) {
    const __attribute__((address_space(4))) uint32_t* tb = (const __attribute__((address_space(4))) uint32_t*)a.tile_batch;
    return batch_at(a.u.batches, tb[tile]);
}

// A lane array of a tile's rows, addressed from the tile's first row: a uniform base (scalar
// registers) and a 32-bit offset per row, so that each load is one instruction with no 64-bit
// address arithmetic per row.  An array the batch lacks reads its stand-in at offset 0 (mask 0).
// Rows are clamped to the tile's last row inside the batch, so every offset stays inside the
// batch's array (and a bitmap's byte offset is the row's >> 3: tiles start on multiples of 8 rows).
struct TileLane {
    const uint8_t* base;
    uint32_t mask;
};
__device__ __forceinline__ TileLane tile_lane(const void* p, bool on, const void* stand_in, int64_t first_byte) {
    TileLane t;
    t.base = on ? (const uint8_t*)p + first_byte : (const uint8_t*)stand_in;
    t.mask = on ? ~0u : 0u;
    return t;
}
__device__ __forceinline__ uint32_t tl_u8(const TileLane& t, uint32_t off) { return *gptr(t.base + (off & t.mask)); }
__device__ __forceinline__ uint32_t tl_u32(const TileLane& t, uint32_t off) {
    return *gptr((const uint32_t*)(t.base + (off & t.mask)));
}
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 tl_u64(const TileLane& t, uint32_t off) {
    const u32x2_t v = *gptr((const u32x2_t*)(t.base + (off & t.mask)));
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 tl_u128(const TileLane& t, uint32_t off) {
    const u32x4_t v = *gptr((const u32x4_t*)(t.base + (off & t.mask)));
    return make_uint4(v.x, v.y, v.z, v.w);
}
// The words of a value lane of `width` bytes (0: none, validity only) at row offset lr.
__device__ __forceinline__ uint4 tl_value(const TileLane& t, int width, uint32_t lr) {
    uint4 x = make_uint4(0u, 0u, 0u, 0u);
    if (width == 16) {
        x = tl_u128(t, lr * 16u);
    } else if (width == 8) {
        const uint2 y = tl_u64(t, lr * 8u);
        x.x = y.x;
        x.y = y.y;
    } else if (width == 4) {
        x.x = tl_u32(t, lr * 4u);
    }
    return x;
}

// Whether row i of a tile takes part: inside its batch, selected, and passing the fused filter.
// sbyte / fbyte: the selection / filter-validity bitmap bytes holding the row's bit; fval: the
// filter lane.
__device__ __forceinline__ bool row_active(const UpdateArgs& u, int64_t i, int64_t n, uint32_t sbyte, uint32_t fbyte,
                                           int32_t fval) {
    const int bit = (int)(i & 7);
    if (i >= n || !((sbyte >> bit) & 1u)) return false;
    return u.filter_col < 0 || (((fbyte >> bit) & 1u) && fval < u.filter_threshold);
}

// What becomes of an active row's key: false for a reserved row (the null key, a key equal to the
// empty marker); else *k = its entry in the hot set (HOT_EMPTY: none, the row becomes a record) and
// *p = its region.  The count pass and the scatter both classify through this function, so their
// record counts agree.  Nothing here branches on the row: the callers predicate what follows.
__device__ __forceinline__ bool key_class(const uint32_t* hidx, const uint64_t* hrow, int HW, uint32_t hbmask, int K,
                                          int pbits, uint64_t key, bool nul, uint32_t* k, uint32_t* p) {
    const uint64_t h = fmix64(key);
    *k = K ? hot_find(hidx, hrow, HW, hbmask, h, key) : HOT_EMPTY;
    *p = (uint32_t)region_of_hash(pbits, h);
    return !nul && key != EMPTY_KEY;
}

// LDS bytes of the count pass: the hot keys, their index, a counter per region.
__host__ __device__ inline size_t count_lds_bytes(int nhot, int hot_slots, int nparts) {
    return (size_t)nhot * 8 + (size_t)hot_slots * 4 + (size_t)nparts * 4;
}

// Phase A, first pass: the records each scatter workgroup will write to each region.  Workgroup
// c = sub * grid_a + w walks every count_split-th of the tiles scatter workgroup w walks (from its
// sub-th on) and classifies their rows as the scatter does, reading only the keys, the selection
// and the filter lane; count[c * nparts + p] = its rows that become records of region p.  Rows of
// the reserved rows (null keys, the empty marker) are folded here.  Its LDS and registers allow two
// workgroups per CU, so that twice the scatter's waves have loads in flight.
template <class PS>
__global__ __launch_bounds__(PA_BLOCK, 8) void part_count_kernel(PartArgs a) {
    constexpr int MR = PA_JIT_RPT;
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int P = a.nparts;
    const int H = a.hot_slots;
    const int K = a.nhot;
    uint64_t* hk = lds;
    uint32_t* hidx = (uint32_t*)(hk + K);
    uint32_t* cnt = hidx + H;
    for (int k = threadIdx.x; k < K; k += blockDim.x) hk[k] = a.hot_keys[k];
    for (int q = threadIdx.x; q < H; q += blockDim.x) hidx[q] = HOT_EMPTY;
    for (int p = threadIdx.x; p < P; p += blockDim.x) cnt[p] = 0;
    __syncthreads();
    const uint32_t hbmask = (uint32_t)(H / HB - 1);
    build_hot_index(hidx, hk, 1, K, hbmask);
    __syncthreads();
    const int RPT = PS::rpt(a);
    const int64_t T = (int64_t)PA_BLOCK * RPT;
    const int pbits = u.t.pbits;
    const bool fast_key = PS::fast_key(a);
    const bool has_filter = u.filter_col >= 0;
    const int64_t G = a.grid_a;
    const int64_t w = blockIdx.x % a.grid_a, sub = blockIdx.x / a.grid_a;
    for (int64_t tile = u.tile_lo + w + G * sub; tile < u.total_tiles; tile += G * a.count_split) {
        const BatchDesc& b = tile_desc(a, tile);
        const int64_t row0 = (tile - b.tile_begin) * T;
        const int64_t left = b.n - row0;  // >= 1: a tile holds rows of its batch
        const uint32_t last = (uint32_t)(left < T ? left : T) - 1u;
        const DCol& fcol = b.vals[has_filter ? u.filter_col : 0];
        const DCol& kcol = b.keys[0];
        const TileLane sel = tile_lane(b.sel, b.sel != nullptr, k_all_ones, row0 >> 3);
        const TileLane fvalid = tile_lane(fcol.valid, has_filter && fcol.valid, k_all_ones, row0 >> 3);
        const TileLane fdata = tile_lane(fcol.data, has_filter, k_zero_words, row0 * 4);
        const TileLane kvalid = tile_lane(kcol.valid, fast_key && kcol.valid, k_all_ones, row0 >> 3);
        const TileLane kdata = tile_lane(kcol.data, fast_key, k_zero_words, row0 * 8);
        uint32_t sbyte[MR], fbyte[MR], kbyte[MR], klo[MR], khi[MR];
        int32_t fval[MR];
#pragma unroll
        for (int r = 0; r < MR; r++) {
            sbyte[r] = fbyte[r] = kbyte[r] = klo[r] = khi[r] = 0;
            fval[r] = 0;
            if (r >= RPT) continue;
            const uint32_t lr = min((uint32_t)(r * PA_BLOCK) + threadIdx.x, last);
            sbyte[r] = tl_u8(sel, lr >> 3);
            fbyte[r] = tl_u8(fvalid, lr >> 3);
            fval[r] = (int32_t)tl_u32(fdata, lr * 4u);
            kbyte[r] = tl_u8(kvalid, lr >> 3);
            const uint2 kw = tl_u64(kdata, lr * 8u);
            klo[r] = kw.x;
            khi[r] = kw.y;
        }
#pragma unroll
        for (int r = 0; r < MR; r++) {
            if (r >= RPT) continue;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            if (!row_active(u, i, b.n, sbyte[r], fbyte[r], fval[r])) continue;
            uint64_t key = 0;
            bool nul;
            if (fast_key) {
                nul = !((kbyte[r] >> (int)(i & 7)) & 1u);
                key = (uint64_t)klo[r] | ((uint64_t)khi[r] << 32);
            } else {
                uint64_t slot;
                nul = !key64_of(u, b, i, &key, &slot);
            }
            uint32_t k, p;
            const bool keyed = key_class(hidx, hk, 1, hbmask, K, pbits, key, nul, &k, &p);
            if (keyed && k == HOT_EMPTY) {
                atomicAdd(&cnt[p], 1u);
            } else if (!keyed) {
                // a reserved row (the null of a 64-bit lane, or a key equal to the empty marker):
                // folded here, with device atomics, so that the scatter's row loop never holds
                // the generic per-row operations
                const uint64_t slot = nul ? u.t.slots + 1 : u.t.slots;
                mark_reserved(u.t.counters, nul ? RES_NULL : RES_SENTINEL);
                apply_ops(u, b, i, u.t.state + slot * (uint64_t)u.t.stride);
            }
        }
        vmem_wait_all();  // the tile's loads are in (see the scatter): the next tile's issue at once
    }
    __syncthreads();
    uint32_t* out = a.count + (size_t)blockIdx.x * P;
    for (int p = threadIdx.x; p < P; p += blockDim.x) out[p] = cnt[p];
}

// base[w * nparts + p] = records of region p that the scatter workgroups before w write (a scan
// down each column of count, the count workgroups of one scatter workgroup summed first, their sum
// left in count[w * nparts + p]); rtot[p] = the column's total.  A workgroup takes 64 regions, each
// of its 16 waves a sixteenth of the scatter workgroups.
__global__ __launch_bounds__(1024) void part_col_scan_kernel(PartArgs a) {
    __shared__ uint32_t part[16][64];
    const int G = a.grid_a, P = a.nparts, CS = a.count_split;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const int per = (G + 15) / 16;
    const int w0 = wv * per < G ? wv * per : G, w1 = w0 + per < G ? w0 + per : G;
    auto total = [&](int w) {
        uint32_t c = 0;
        for (int q = 0; q < CS; q++) c += a.count[((size_t)q * G + w) * P + p];
        return c;
    };
    uint32_t s = 0;
    if (p < P)
        for (int w = w0; w < w1; w++) s += total(w);
    part[wv][lane] = s;
    __syncthreads();
    uint32_t off = 0;
    for (int j = 0; j < wv; j++) off += part[j][lane];
    if (p >= P) return;
    for (int w = w0; w < w1; w++) {
        const uint32_t c = total(w);
        if (CS > 1) a.count[(size_t)w * P + p] = c;
        a.base[(size_t)w * P + p] = off;
        off += c;
    }
    if (wv == 15) a.rtot[p] = off;
}

// rstart[p] = where region p's records start: its range is padded to whole blocks of PART_BLK
// records, so that no block of the pool holds two regions' records (rstart[nparts]: the end).
__global__ __launch_bounds__(1024) void part_region_scan_kernel(PartArgs a) {
    __shared__ uint32_t ws[1024];
    const int P = a.nparts;
    const int per = (P + 1023) / 1024;
    const int p0 = threadIdx.x * per;
    auto padded = [&](int q) -> uint32_t { return (a.rtot[q] + PART_BLK - 1) / PART_BLK * PART_BLK; };
    uint32_t s = 0;
    for (int j = 0; j < per; j++)
        if (p0 + j < P) s += padded(p0 + j);
    ws[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint32_t x = threadIdx.x >= (unsigned)d ? ws[threadIdx.x - d] : 0;
        __syncthreads();
        ws[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t off = ws[threadIdx.x] - s;
    for (int j = 0; j < per; j++) {
        if (p0 + j >= P) break;
        a.rstart[p0 + j] = off;
        off += padded(p0 + j);
    }
    if (threadIdx.x == 1023) a.rstart[P] = ws[1023];
}

// LDS bytes of the scatter: the hot rows, their index, a record cursor per region.
__host__ __device__ inline size_t scatter_lds_bytes(int nhot, int hot_slots, int stride, int nparts) {
    return (size_t)nhot * lds_row_words(stride) * 8 + (size_t)hot_slots * 4 + (size_t)nparts * 4;
}

// Word o of a record held in registers (o uniform): a select per word, no indexed register access.
template <int RW>
__device__ __forceinline__ void put_word(uint32_t (&w)[RW], int o, uint32_t x) {
#pragma unroll
    for (int j = 0; j < RW; j++)
        if (j == o) w[j] = x;
}

// Phase A, second pass.  RW: 32-bit words per record (the layout's `words`); PS: where the column
// programs come from.  The workgroup's record cursor of region p starts at rstart[p] + base[w][p],
// so the records of a region are contiguous in the pool (workgroup by workgroup) and a row's
// record position is one LDS atomic away: no wave ever waits for another inside the tile loop.
template <int RW, class PS>
__global__ __launch_bounds__(PA_BLOCK) void part_scatter_kernel(PartArgs a) {
    constexpr int MR = PS::max_rpt;  // rows per thread per tile at most (registers per row)
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int S = u.t.stride;
    const int P = a.nparts;
    const int H = a.hot_slots;
    const int K = a.nhot;
    // LDS: hot keys and states [K][HW], hot index [H] (u32), record cursor per region [P] (u32)
    const int HW = lds_row_words(S);
    uint64_t* hrow = lds;  // row k: key k of the hot set, its state words
    uint32_t* hidx = (uint32_t*)(hrow + (size_t)K * HW);
    uint32_t* cursor = hidx + H;
    const uint32_t* wbase = a.base + (size_t)blockIdx.x * P;
    for (int w = threadIdx.x; w < K * HW; w += blockDim.x) {
        const int j = w % HW;
        hrow[w] = j == 0 ? a.hot_keys[w / HW] : (j <= S ? u.ident[j - 1] : 0);
    }
    for (int q = threadIdx.x; q < H; q += blockDim.x) hidx[q] = HOT_EMPTY;
    for (int p = threadIdx.x; p < P; p += blockDim.x) cursor[p] = a.rstart[p] + wbase[p];
    __syncthreads();
    const uint32_t hbmask = (uint32_t)(H / HB - 1);
    build_hot_index(hidx, hrow, HW, K, hbmask);
    __syncthreads();

    const int RPT = PS::rpt(a);
    const int64_t T = (int64_t)PA_BLOCK * RPT;
    const int pbits = u.t.pbits;
    const bool fast_key = PS::fast_key(a);
    const bool has_filter = u.filter_col >= 0;
    // A tile's records are built in registers and stored at the start of the next tile, before
    // its loads: the loads then never wait behind the stores (memory operations complete in
    // order), and a store's registers are written again only after the next tile's loads arrived,
    // by when the store has read them.  Records wider than 32 bytes are stored at once (their
    // registers would not fit).
    constexpr bool kDelay = RW <= 8;
    uint32_t rec[MR][RW];
    uint32_t rec_at[MR];  // the pool position of rec[r]
    uint32_t rec_on = 0;  // bit r: rec[r] holds a record to store
#pragma unroll
    for (int r = 0; r < MR; r++) rec_at[r] = 0;
    auto store_record = [&](int r) __attribute__((always_inline)) {
        if (!((rec_on >> r) & 1u)) return;
        rec_on &= ~(1u << r);
        uint4* dst = (uint4*)(a.pool + (uint64_t)rec_at[r] * RW);
#pragma unroll
        for (int q = 0; q < RW / 4; q++)
            dst[q] = make_uint4(rec[r][4 * q], rec[r][4 * q + 1], rec[r][4 * q + 2], rec[r][4 * q + 3]);
    };
    // Tiles are grid-strided (the CUs read neighbouring addresses at any time), the same tiles per
    // workgroup as the count pass.
    for (int64_t tile = u.tile_lo + blockIdx.x; tile < u.total_tiles; tile += gridDim.x) {
        const BatchDesc& b = tile_desc(a, tile);
        const int64_t row0 = (tile - b.tile_begin) * T;
        if (kDelay) {
#pragma unroll
            for (int r = 0; r < MR; r++) store_record(r);
        }
        // Every load of the tile first, with no branch between them (one round trip per tile): the
        // row index is clamped into the batch, a lane is read as 32-bit words at i * width (the
        // second word re-reads the first for a 4-byte lane), bitmaps the batch lacks are read from
        // k_all_ones.  Validity, selection and the filter are applied afterwards.
        uint32_t vbyte[MR][PART_MAX_PROG];
        uint32_t lo[MR][PART_MAX_PROG], hi[MR][PART_MAX_PROG];
        uint32_t w2[MR][PART_MAX_PROG], w3[MR][PART_MAX_PROG];
        uint32_t sbyte[MR], fbyte[MR], kbyte[MR];
        int32_t fval[MR];
        uint32_t klo[MR], khi[MR];
        const DCol& fcol = b.vals[has_filter ? u.filter_col : 0];
        const DCol& kcol = b.keys[0];
        const int64_t left = b.n - row0;  // >= 1: a tile holds rows of its batch
        const uint32_t last = (uint32_t)(left < T ? left : T) - 1u;
        const TileLane sel = tile_lane(b.sel, b.sel != nullptr, k_all_ones, row0 >> 3);
        const TileLane fvalid = tile_lane(fcol.valid, has_filter && fcol.valid, k_all_ones, row0 >> 3);
        const TileLane fdata = tile_lane(fcol.data, has_filter, k_zero_words, row0 * 4);
        // the 64-bit key lane (other key modes read their lanes while classifying)
        const TileLane kvalid = tile_lane(kcol.valid, fast_key && kcol.valid, k_all_ones, row0 >> 3);
        const TileLane kdata = tile_lane(kcol.data, fast_key, k_zero_words, row0 * 8);
        TileLane vvalid[PART_MAX_PROG], vdata[PART_MAX_PROG];
        int vwidth[PART_MAX_PROG];
#pragma unroll
        for (int q = 0; q < PART_MAX_PROG; q++) {
            const bool on = q < PS::nprog(a);
            const ColProg& cp = PS::prog(a, on ? q : 0);
            const DCol& d = b.vals[cp.col];
            vwidth[q] = on ? cp.width : 0;
            vvalid[q] = tile_lane(d.valid, on && d.valid, k_all_ones, row0 >> 3);
            vdata[q] = tile_lane(d.data, vwidth[q] != 0, k_zero_words, row0 * vwidth[q]);
        }
#pragma unroll
        for (int r = 0; r < MR; r++) {
            const uint32_t lr = min((uint32_t)(r * PA_BLOCK) + threadIdx.x, last);
            sbyte[r] = tl_u8(sel, lr >> 3);
            fbyte[r] = tl_u8(fvalid, lr >> 3);
            fval[r] = (int32_t)tl_u32(fdata, lr * 4u);
            kbyte[r] = tl_u8(kvalid, lr >> 3);
            const uint2 kw = tl_u64(kdata, lr * 8u);
            klo[r] = kw.x;
            khi[r] = kw.y;
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++) {
                vbyte[r][q] = tl_u8(vvalid[q], lr >> 3);
                const uint4 x = tl_value(vdata[q], vwidth[q], lr);
                lo[r][q] = x.x;
                hi[r][q] = x.y;
                w2[r][q] = x.z;
                w3[r][q] = x.w;
            }
        }
        // The rows' fates, with no branch on the row: hot entry (hot_m), record (rec_m), neither.
        // fate[r]: the hot entry << 16 | the region (both below 2^16); vm: 4 validity bits per row.
        uint32_t fate[MR];
        uint32_t hot_m = 0, rec_m = 0, vm = 0;
#pragma unroll
        for (int r = 0; r < MR; r++) {
            fate[r] = 0;
            if (r >= RPT) break;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            const bool act = row_active(u, i, b.n, sbyte[r], fbyte[r], fval[r]);
            const int bit = (int)(i & 7);
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++)
                if (q < PS::nprog(a) && ((vbyte[r][q] >> bit) & 1u)) vm |= 1u << (4 * r + q);
            uint64_t key = 0;
            bool nul;
            if (fast_key) {
                nul = !((kbyte[r] >> bit) & 1u);
                key = (uint64_t)klo[r] | ((uint64_t)khi[r] << 32);
            } else {
                uint64_t slot;
                nul = act ? !key64_of(u, b, i, &key, &slot) : true;
            }
            uint32_t k, p;
            const bool keyed = key_class(hidx, hrow, HW, hbmask, K, pbits, key, nul, &k, &p) && act;
            // rows of the reserved rows were folded by the count pass
            hot_m |= (uint32_t)(keyed && k != HOT_EMPTY) << r;
            rec_m |= (uint32_t)(keyed && k == HOT_EMPTY) << r;
            fate[r] = (k << 16) | p;
        }
        // record positions: the LDS atomics of every record row issued together
        uint32_t pos[MR];
#pragma unroll
        for (int r = 0; r < MR; r++) {
            pos[r] = 0;
            if ((rec_m >> r) & 1u) pos[r] = atomicAdd(&cursor[fate[r] & 0xFFFFu], 1u);
        }
        // hot rows: COUNT(*), then each column's program, into the workgroup's LDS copy
#pragma unroll
        for (int r = 0; r < MR; r++) {
            if (!((hot_m >> r) & 1u)) continue;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            uint64_t* e = hrow + (size_t)(fate[r] >> 16) * HW + 1;
            atomic_add_u64(e, 1);
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++) {
                if (q >= PS::nprog(a)) break;
                const ColProg& cp = PS::prog(a, q);
                const int wd = cp.width;
                const uint64_t xa = (uint64_t)lo[r][q] | (wd >= 8 ? (uint64_t)hi[r][q] << 32 : 0ull);
                const uint64_t xb = wd >= 16 ? ((uint64_t)w2[r][q] | ((uint64_t)w3[r][q] << 32)) : 0ull;
                const bool v = (vm >> (4 * r + q)) & 1u;
                if (cp.flags & CP_GENERIC) {
                    const RegSrc src{v, xa, xb, (uint64_t)(b.row_base + i), cp.type};
                    for (int k = 0; k < u.nops; k++)
                        if (u.ops[k].col == cp.col) apply_one(u, u.ops[k], src, e);
                } else {
                    apply_prog(u, cp, v, xa, xb, e);
                }
            }
        }
        // records: key, the carried lanes, validity, row index, stored at the region's cursor
#pragma unroll
        for (int r = 0; r < MR; r++) {
            if (!((rec_m >> r) & 1u)) continue;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            uint32_t(&w)[RW] = rec[r];
#pragma unroll
            for (int j = 0; j < RW; j++) w[j] = 0;
            w[0] = klo[r];
            w[1] = khi[r];
            if (!fast_key) {
                uint64_t key = 0, slot;
                (void)key64_of(u, b, i, &key, &slot);
                w[0] = (uint32_t)key;
                w[1] = (uint32_t)(key >> 32);
            }
            uint32_t vbits = 0;
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++) {
                if (q >= PS::nprog(a)) break;
                const ColProg& cp = PS::prog(a, q);
                if ((vm >> (4 * r + q)) & 1u) vbits |= 1u << cp.col;
                const int o = cp.rec_off;
                if (o < 0) continue;
                put_word<RW>(w, o, lo[r][q]);
                if (cp.type != VR_INT32) put_word<RW>(w, o + 1, hi[r][q]);
                if (cp.type == VR_DECIMAL128) {
                    put_word<RW>(w, o + 2, w2[r][q]);
                    put_word<RW>(w, o + 3, w3[r][q]);
                }
            }
            put_word<RW>(w, PS::valid_off(a), vbits);
            if (PS::row_off(a) >= 0) {
                const uint64_t gi = (uint64_t)(b.row_base + i);
                put_word<RW>(w, PS::row_off(a), (uint32_t)gi);
                put_word<RW>(w, PS::row_off(a) + 1, (uint32_t)(gi >> 32));
            }
            // a position past the pool (the count pass disagreeing, never expected) is kept in the
            // pool; the range check after the loop fails the launch
            rec_at[r] = pos[r] < a.pool_records ? pos[r] : a.pool_records - 1;
            rec_on |= 1u << r;
            if (!kDelay) store_record(r);
        }
        // Every load of the tile has arrived by now (the stores issued before them are done too):
        // saying so here keeps the compiler from holding the next tile's first instructions for
        // the loads of rows that took no part, whose values were never read.
        if (kDelay) vmem_wait_all();
    }
    if (kDelay)
        for (int r = 0; r < MR; r++) store_record(r);
    __syncthreads();
    // every range filled exactly (the count pass classified the same rows), and the hot partials
    const uint32_t* wcount = a.count + (size_t)blockIdx.x * P;
    for (int p = threadIdx.x; p < P; p += blockDim.x)
        if (cursor[p] != a.rstart[p] + wbase[p] + wcount[p])
            atomicOr((unsigned long long*)&u.t.counters[C_ERROR], (unsigned long long)ERR_INTERNAL);
    uint64_t* hp = a.hot_part + (size_t)blockIdx.x * K * S;
    for (int w = threadIdx.x; w < K * S; w += blockDim.x) hp[w] = hrow[(size_t)(w / S) * HW + 1 + w % S];
}

// hot_total[k] = the hot partials of every scatter workgroup combined.
__global__ __launch_bounds__(256) void hot_reduce_kernel(PartArgs a) {
    const int S = a.u.t.stride;
    const int KS = a.nhot * S;
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < KS; w += gridDim.x * blockDim.x) {
        const int k = w % S;
        const uint8_t code = a.u.comb[k];
        uint64_t acc = a.u.ident[k];
        for (int g = 0; g < a.grid_a; g++) acc = combine_word(code, acc, a.hot_part[(size_t)g * KS + w]);
        a.hot_total[w] = acc;
    }
}

// ------------------------------------------------------------------ phase B: merge

// LDS bytes of the merge: the cache (E keys and state rows).
__host__ __device__ inline size_t merge_lds_bytes(int E, int stride) {
    return (size_t)E * lds_row_words(stride) * 8;
}

// Entry of `key` in the LDS cache, claiming an empty one; -1 when neither is found within `limit`
// buckets (limit 0: no limit -- the caller guarantees a free entry).  The cache is a table of
// buckets of CB keys, filled from each bucket's start: one bucket read answers most lookups.
constexpr int CB = 8;
__device__ __forceinline__ int cache_find(uint64_t* crow, int CW, int E, uint64_t h, uint64_t key, int limit) {
    const uint32_t bmask = (uint32_t)(E / CB - 1);
    uint32_t b = (uint32_t)h & bmask;
    for (int q = 0; limit == 0 || q < limit;) {
        const uint64_t* p = crow + (size_t)b * CB * CW;
        uint64_t k[CB];
#pragma unroll
        for (int j = 0; j < CB; j++) k[j] = p[(size_t)j * CW];
        int empty = -1;
#pragma unroll
        for (int j = CB - 1; j >= 0; j--) {
            if (k[j] == key) return (int)(b * CB + j);
            if (k[j] == EMPTY_KEY) empty = j;
        }
        if (empty < 0) {
            b = (b + 1) & bmask;
            q++;
            continue;
        }
        const uint64_t old = atomicCAS((unsigned long long*)&crow[(size_t)(b * CB + empty) * CW],
                                       (unsigned long long)EMPTY_KEY, (unsigned long long)key);
        if (old == EMPTY_KEY || old == key) return (int)(b * CB + empty);
        // another lane took the entry: read the bucket again
    }
    return -1;
}

// Folds a record held in registers into a cache entry: COUNT(*), then each column's program.
template <int RW, class PS>
__device__ __forceinline__ void apply_record(const PartArgs& a, const uint32_t (&w)[RW], uint64_t* e) {
    const UpdateArgs& u = a.u;
    atomic_add_u64(e, 1);
    const uint32_t vb = get_word<RW>(w, PS::valid_off(a));
    for (int q = 0; q < PS::nprog(a); q++) {
        const ColProg cp = PS::prog(a, q);
        const bool v = (vb >> cp.col) & 1u;
        uint64_t x0 = 0, x1 = 0;
        const int o = cp.rec_off;
        if (o >= 0 && v) {
            x0 = get_word<RW>(w, o);
            if (cp.type != VR_INT32) x0 |= (uint64_t)get_word<RW>(w, o + 1) << 32;
            if (cp.type == VR_DECIMAL128) x1 = (uint64_t)get_word<RW>(w, o + 2) | ((uint64_t)get_word<RW>(w, o + 3) << 32);
        }
        if (cp.flags & CP_GENERIC) {
            const int ro = PS::row_off(a);
            const uint64_t gi = ro >= 0 ? ((uint64_t)get_word<RW>(w, ro) | ((uint64_t)get_word<RW>(w, ro + 1) << 32))
                                    : 0;
            const RegSrc src{v, x0, x1, gi, cp.type};
            for (int k = 0; k < u.nops; k++)
                if (u.ops[k].col == cp.col) apply_one(u, u.ops[k], src, e);
        } else {
            apply_prog(u, cp, v, x0, x1, e);
        }
    }
}

template <int RW, class PS>
__global__ __launch_bounds__(PB_BLOCK) void part_merge_kernel(PartArgs a) {
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int S = u.t.stride;
    const int E = a.cache_e;
    const uint32_t p = blockIdx.x;
    const int CW = lds_row_words(S);
    uint64_t* crow = lds;  // cache entry e: row e (key, state words)
    const int lane = threadIdx.x & 63;
    __shared__ unsigned int s_count, s_inserted, s_rg, s_new, s_more;
    TableView tv = u.t;
    tv.rgroups = nullptr;  // inserts are counted here, once per merge
    const uint32_t limit = a.region_limit;

    int resume = a.resumed ? a.resume_step[p] : 0;
    if (resume < 0) return;  // merged already

    if (threadIdx.x == 0) {
        s_count = 0;
        s_inserted = 0;
        s_new = 0;
        s_rg = u.t.rgroups[p];  // the region is this workgroup's alone: its group count stays here
        s_more = 0;
    }
    __syncthreads();
    // after a barrier that follows the inserts; read by every thread after the next barrier
    auto add_inserted = [&]() {
        if (threadIdx.x == 0) {
            s_rg += s_inserted;
            s_new += s_inserted;
            s_inserted = 0;
            s_count = 0;
            s_more = 0;
        }
    };
    bool stopped = false;
    if (resume == 0) {
        // the hot set's partials of this region, straight into the region (one thread per key)
        for (int k = threadIdx.x; k < a.nhot; k += blockDim.x) {
            const uint64_t key = a.hot_keys[k];
            if (key == EMPTY_KEY || a.hot_total[(size_t)k * S] == 0) continue;
            if (region_of_hash(u.t.pbits, fmix64(key)) == p) atomicAdd(&s_count, 1u);
        }
        __syncthreads();
        if (s_rg + s_count > limit) {
            stopped = true;
        } else {
            for (int k = threadIdx.x; k < a.nhot; k += blockDim.x) {
                const uint64_t key = a.hot_keys[k];
                if (key == EMPTY_KEY || a.hot_total[(size_t)k * S] == 0) continue;
                if (region_of_hash(u.t.pbits, fmix64(key)) != p) continue;
                bool ins;
                const uint64_t slot = find_or_insert64(tv, key, &ins);
                if (ins) atomicAdd(&s_inserted, 1u);
                merge_plain(u, a.hot_total + (size_t)k * S, u.t.state + slot * (uint64_t)S, S);
            }
            __syncthreads();
            add_inserted();
            resume = 1;
        }
    }

    auto reset_cache = [&]() {
        for (int w = threadIdx.x; w < E * CW; w += blockDim.x) {
            const int j = w % CW;
            crow[w] = j == 0 ? EMPTY_KEY : (j <= S ? u.ident[j - 1] : 0);
        }
    };
    reset_cache();
    __syncthreads();

    // The new slots this merge claims: the region is this workgroup's alone, so an empty cell it
    // reads can only be taken by another of its threads, and a bit per slot in LDS settles which
    // (no device atomic).  A later pass of the workgroup reads the cells after barriers, which
    // order its own stores.  Regions of more than PB_CLAIM_BITS slots claim by CAS instead.
    __shared__ uint32_t s_claim[PB_CLAIM_BITS / 32];
    const bool lds_claim = tv.rmask < (uint64_t)PB_CLAIM_BITS;

    // Merges every cache entry that took rows into the region.  Returns false, merging nothing,
    // when the region could not take every such entry as a new group.  Called by every thread.
    // *more: whether a record of the pass found no cache entry (read here by every thread).
    auto flush = [&](bool* more) __attribute__((always_inline)) -> bool {
        __syncthreads();  // every fold before the call is done
        for (int w = threadIdx.x; w < PB_CLAIM_BITS / 32; w += blockDim.x) s_claim[w] = 0;
        unsigned int mine = 0;
        for (int e = threadIdx.x; e < E; e += blockDim.x)
            if (crow[(size_t)e * CW] != EMPTY_KEY && crow[(size_t)e * CW + 1] != 0) mine++;
        if (mine) atomicAdd(&s_count, mine);
        __syncthreads();
        *more = s_more != 0;
        // uniform: s_rg, s_count and s_more change after the next barrier
        if (s_rg + s_count > limit) return false;
        // this thread's PB_FE entries, side by side: every global access of one is issued with
        // the other's (the entries' slots, then their state rows)
        uint64_t fk[PB_FE], slot[PB_FE], rows[PB_FE];
        const uint64_t* l[PB_FE];
#pragma unroll
        for (int j = 0; j < PB_FE; j++) {
            const int e = threadIdx.x + j * PB_BLOCK;
            fk[j] = crow[(size_t)e * CW];
            l[j] = crow + (size_t)e * CW + 1;
            rows[j] = fk[j] == EMPTY_KEY ? 0 : l[j][0];
            slot[j] = home_slot(tv, fmix64(fk[j]));
        }
        uint64_t cell[PB_FE];
#pragma unroll
        for (int j = 0; j < PB_FE; j++) cell[j] = tv.cells[slot[j]];
        unsigned int ins = 0;
        bool fresh[PB_FE];  // a new group: its state row is at identity, no need to read it
#pragma unroll
        for (int j = 0; j < PB_FE; j++) {
            fresh[j] = false;
            if (!rows[j]) continue;
            for (uint64_t probes = 0;; probes++) {  // linear probing inside the region, as find_or_insert64
                if (cell[j] == fk[j]) break;
                if (cell[j] == EMPTY_KEY) {
                    bool won;
                    if (lds_claim) {
                        const uint32_t r = (uint32_t)(slot[j] & tv.rmask), bit = 1u << (r & 31);
                        won = (atomicOr(&s_claim[r >> 5], bit) & bit) == 0;
                        if (won) tv.cells[slot[j]] = fk[j];
                    } else {
                        const uint64_t old = atomic_cas_u64(&tv.cells[slot[j]], EMPTY_KEY, fk[j]);
                        won = old == EMPTY_KEY;
                        if (old == fk[j]) break;
                    }
                    if (won) {
                        ins++;
                        fresh[j] = true;
                        break;
                    }
                }
                if (probes == tv.rmask) {
                    slot[j] = region_full(tv);
                    break;
                }
                slot[j] = next_slot(tv, slot[j]);
                cell[j] = tv.cells[slot[j]];
            }
        }
        if (ins) atomicAdd(&s_inserted, ins);
        for (int k0 = 0; k0 < S; k0 += 8) {
            uint64_t y[PB_FE][8];
#pragma unroll
            for (int j = 0; j < PB_FE; j++) {
                const uint64_t* g = u.t.state + slot[j] * (uint64_t)S;
#pragma unroll
                for (int q = 0; q < 8; q++) y[j][q] = fresh[j] ? 0 : g[k0 + q < S ? k0 + q : S - 1];
            }
#pragma unroll
            for (int j = 0; j < PB_FE; j++) {
                if (!rows[j]) continue;
                uint64_t* g = u.t.state + slot[j] * (uint64_t)S;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int k = k0 + q;
                    if (k >= S) break;
                    const uint64_t x = l[j][k];
                    if (fresh[j]) g[k] = x;  // identity combined with x
                    else if (x != u.ident[k]) g[k] = combine_word(u.comb[k], y[j][q], x);
                }
            }
        }
        __syncthreads();
        add_inserted();
        reset_cache();
        __syncthreads();
        return true;
    };

    // The region's blocks (its records are contiguous from rstart[p], PART_BLK to a block, the
    // last one partly filled), handed out to the waves as they ask with an LDS counter: the waves'
    // work evens out whatever the records' keys.
    __shared__ unsigned int s_gnext;
    const uint32_t rblock0 = a.rstart[p] / PART_BLK;
    uint32_t rrecs = a.rtot[p];
    uint32_t rblocks = (rrecs + PART_BLK - 1) / PART_BLK;
    if ((uint64_t)rblock0 + rblocks > a.pool_blocks) {
        // ranges past the pool (the counts disagreeing with the pool's bound, never expected): the
        // blocks inside it are read, and the launch fails
        if (threadIdx.x == 0) atomicOr((unsigned long long*)&a.u.t.counters[C_ERROR], (unsigned long long)ERR_INTERNAL);
        rblocks = rblock0 < a.pool_blocks ? (uint32_t)(a.pool_blocks - rblock0) : 0u;
        rrecs = rblocks * PART_BLK;
    }
    // the next block for this wave (NO_BLOCK: none) and its record count
    auto fetch = [&](uint32_t* n) -> uint32_t {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&s_gnext, 1u);
        i = __shfl(i, 0, 64);
        if (i >= rblocks) {
            *n = 0;
            return NO_BLOCK;
        }
        *n = i + 1 < rblocks ? (uint32_t)PART_BLK : rrecs - i * PART_BLK;
        return rblock0 + i;
    };

    // Passes.  Pass 0 folds every record; a record that finds no cache entry (the cache is full)
    // is marked in its block's word of the other pending map, and pass k >= 1 folds the records
    // marked by pass k - 1.  Each pass ends with one merge of the cache into the region.  The two
    // maps are used in turn, so a pass whose merge finds the region short stops with its input
    // intact and resumes there after the table grows.  Within a pass the waves run on their own
    // (no barrier): each loads the records of its blocks two blocks ahead, into three buffers
    // used in turn (each unrolled step names its buffers, so no register copy waits for a load in
    // flight).
    int pass = resume - 1;
    uint32_t ra[RW], rb[RW], rc[RW];
    while (!stopped) {
        const uint64_t* rd = pass ? a.pending + (size_t)(pass & 1) * a.pool_blocks : nullptr;
        uint64_t* wr = a.pending + (size_t)((pass + 1) & 1) * a.pool_blocks;
        if (threadIdx.x == 0) s_gnext = 0;
        __syncthreads();
        // this lane's record of block b (n records); true when it is pending in this pass
        auto load = [&](uint32_t b, uint32_t n, uint32_t (&w)[RW]) -> bool {
            if (b == NO_BLOCK || (uint32_t)lane >= n) return false;
            const uint4* src = (const uint4*)gptr(a.pool + ((uint64_t)b * PART_BLK + lane) * (uint64_t)RW);
#pragma unroll
            for (int q = 0; q < RW / 4; q++) {
                const uint4 x = src[q];
                w[4 * q] = x.x;
                w[4 * q + 1] = x.y;
                w[4 * q + 2] = x.z;
                w[4 * q + 3] = x.w;
            }
            return rd == nullptr || ((*gptr(rd + b) >> lane) & 1ull);
        };
        // folds buffer X (block bX) and marks its records that found no entry; loads the block
        // after next into Z.  False when the wave has no block after X.
        auto run_step = [&](const uint32_t (&X)[RW], bool hX, uint32_t bX, uint32_t bY, uint32_t (&Z)[RW],
                            uint32_t& bZ, bool& hZ) __attribute__((always_inline)) -> bool {
            uint32_t nz;
            bZ = fetch(&nz);
            hZ = load(bZ, nz, Z);
            int e = -1;
            if (hX) {
                const uint64_t key = (uint64_t)X[0] | ((uint64_t)X[1] << 32);
                e = cache_find(crow, CW, E, fmix64(key), key, E / CB);  // every bucket if need be
                if (e >= 0) apply_record<RW, PS>(a, X, crow + (size_t)e * CW + 1);
            }
            const uint64_t missed = __ballot(hX && e < 0);
            if (bX != NO_BLOCK && lane == 0) {
                wr[bX] = missed;
                if (missed) s_more = 1;
            }
            return bY != NO_BLOCK;
        };
        {
            uint32_t ba, bb, bc = NO_BLOCK, n;
            bool ha, hb, hc = false;
            ba = fetch(&n);
            ha = load(ba, n, ra);
            bb = fetch(&n);
            hb = load(bb, n, rb);
            if (ba != NO_BLOCK)
                while (run_step(ra, ha, ba, bb, rc, bc, hc) && run_step(rb, hb, bb, bc, ra, ba, ha) &&
                       run_step(rc, hc, bc, ba, rb, bb, hb)) {
                }
        }
        bool more;
        if (!flush(&more)) {
            stopped = true;  // this pass resumes after the table grows
            break;
        }
        if (!more) break;
        pass++;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u.t.rgroups[p] = s_rg;
        if (s_new) atomic_add_u64(&u.t.counters[C_GROUPS], s_new);
        // -1: merged; 0: the hot partials too; k + 1: pass k
        a.resume_step[p] = stopped ? (resume == 0 ? 0 : pass + 1) : -1;
        if (stopped) atomicOr((unsigned long long*)&u.t.counters[C_PSTOP], 1ull);
    }
}

}  // namespace vr
