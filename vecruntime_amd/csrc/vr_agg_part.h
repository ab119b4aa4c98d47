// Partitioned group-by update for keys of at most 64 bits: the batched path without per-row
// device atomics.
//
// Device-scope atomics on this part run at ~24 G/s whatever the working set, while LDS atomics run
// at ~2 T/s and a wave streams a block of 64 records at HBM speed.  The update is therefore split
// into two launches in which no global state row is ever updated by two workgroups:
//
//   A (scatter)  Each workgroup streams tiles of rows once.  Rows whose key is in the hot set (the
//                most frequent keys of a sample of the input) fold into a per-workgroup LDS copy
//                of those groups.  Every other row becomes a record -- the key, the value lanes the
//                aggregates read, the row index when a double MIN/MAX needs it, one validity word
//                -- appended to its region's chain of 64-record blocks: a workgroup keeps one open
//                block per region, takes new blocks from a pool with one atomic per block and
//                links each to the previous one (block_next), so the records of region r are the
//                chains heads[w][r] of all scatter workgroups w.
//   B (merge)    One workgroup per region owns that region's slots and state rows.  It first
//                merges the hot set's partials (reduced over the scatter workgroups) of its region,
//                then its waves walk the chains in lockstep steps of one block each, folding
//                records into an LDS cache of groups with LDS atomics.  When a step's record finds
//                no room in the cache, the cache is merged into the region -- new slots claimed by
//                CAS, state rows by plain read-modify-write, since no other workgroup touches the
//                region -- keeping the entries that took two rows or more.
//
// Growth.  A merge first checks that the region can take every cached group as a new one; when
// it cannot, the workgroup stops and leaves in resume_step the first step (+1; 0: the hot
// partials too) not merged yet.  The table grows (regions double, their number is kept) and the
// stopped regions resume there.
#pragma once

#include "vr_agg_kernels.h"

namespace vr {

constexpr int PART_BLK = 64;         // records per block: one per lane of a wave
constexpr int PART_REC_MAX = 32;     // 32-bit words per record
constexpr uint32_t NO_BLOCK = 0xFFFFFFFFu;
constexpr int PA_BLOCK = 512;        // scatter workgroup
constexpr int PA_MAX_RPT = 8;        // rows per thread per scatter tile
constexpr int PB_BLOCK = 512;        // merge workgroup
constexpr int PB_PROBES = 16;        // LDS cache probes before a record waits for the next merge
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

struct PartArgs {
    UpdateArgs u;
    RecLayout rl;
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
    uint32_t* block_next;      // the chain's previous block
    uint32_t* block_cnt;       // records in the block
    uint32_t* heads;           // [grid_a][nparts] newest block of each chain
    int32_t* resume_step;      // [nparts] merge: see above; -1 when merged
    uint64_t max_blocks;
};

// A record in the pool, read field by field by the merge.
struct RecSrc {
    const uint32_t* r;
    const RecLayout& L;
    __device__ __forceinline__ bool valid(int c) const { return (r[L.valid_off] >> c) & 1u; }
    __device__ __forceinline__ uint64_t u64(int c) const { return *(const uint64_t*)(r + L.val_off[c]); }
    __device__ __forceinline__ int32_t i32(int c) const { return (int32_t)r[L.val_off[c]]; }
    __device__ __forceinline__ int64_t lng(int c) const {
        return L.val_type[c] == VR_INT32 ? (int64_t)i32(c) : (int64_t)u64(c);
    }
    __device__ __forceinline__ void d128(int c, uint64_t* lo, uint64_t* hi) const {
        const uint64_t* p = (const uint64_t*)(r + L.val_off[c]);
        *lo = p[0];
        *hi = p[1];
    }
    __device__ __forceinline__ uint64_t row() const { return *(const uint64_t*)(r + L.row_off); }
};

// Writes the record of row i to `dst` (8-byte aligned): key, carried value lanes, row index,
// validity bits.  Lanes of null values are left unwritten (never read).
__device__ __forceinline__ void write_record(const RecLayout& L, const BatchDesc& b, int64_t i, uint64_t key,
                                             uint32_t* dst) {
    *(uint64_t*)dst = key;
    uint32_t vbits = 0;
    for (int c = 0; c < L.nvals; c++) {
        if (!((L.vmask >> c) & 1u)) continue;
        const DCol& col = b.vals[c];
        if (!is_valid(col, i)) continue;
        vbits |= 1u << c;
        const int o = L.val_off[c];
        if (o < 0) continue;
        switch (L.val_type[c]) {
            case VR_INT32: dst[o] = (uint32_t)ld_stream((const int32_t*)col.data + i); break;
            case VR_INT64:
            case VR_FLOAT64: *(uint64_t*)(dst + o) = (uint64_t)ld_stream((const int64_t*)col.data + i); break;
            case VR_DECIMAL128: {
                const uint64_t* p = (const uint64_t*)col.data + 2 * i;
                uint64_t lo = ld_stream(p), hi = ld_stream(p + 1);
                *(uint64_t*)(dst + o) = lo;
                *(uint64_t*)(dst + o + 2) = hi;
                break;
            }
            default: break;
        }
    }
    if (L.row_off >= 0) *(uint64_t*)(dst + L.row_off) = (uint64_t)(b.row_base + i);
    dst[L.valid_off] = vbits;
}

__device__ __forceinline__ const BatchDesc& batch_of_tile(const UpdateArgs& u, int64_t tile) {
    if (u.num_batches <= 1) return u.one;
    int lo = 0, hi = u.num_batches - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (u.batches[mid].tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return u.batches[lo];
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

// ------------------------------------------------------------------ hot set

// Counts the keys of `per` consecutive rows at the start of `nsamp` evenly spaced tiles (tiles of
// PA_BLOCK * rows_per_thread rows) in `cand`: open addressing over cand_mask + 1 entries of (key,
// count).  Keys past a full candidate table are dropped (the hot set is only a hint).
__global__ __launch_bounds__(256) void hot_sample_kernel(UpdateArgs u, int per, int64_t nsamp, uint64_t* cand,
                                                         uint32_t* cand_count, uint64_t cand_mask) {
    const int64_t T = (int64_t)u.rows_per_thread * PA_BLOCK;
    for (int64_t s = blockIdx.x; s < nsamp; s += gridDim.x) {
        const int64_t tile = u.tile_lo + s * ((u.total_tiles - u.tile_lo) / nsamp);
        const BatchDesc& b = batch_of_tile(u, tile);
        const int64_t row0 = (tile - b.tile_begin) * T;
        for (int r = threadIdx.x; r < per; r += blockDim.x) {
            const int64_t i = row0 + r;
            if (i >= b.n) break;
            if (b.sel && !bit_at(b.sel, i)) continue;
            uint64_t key, slot;
            if (!key64_of(u, b, i, &key, &slot) || key == EMPTY_KEY) continue;
            uint64_t pos = fmix64(key) & cand_mask;
            for (uint64_t q = 0; q <= cand_mask; q++) {
                uint64_t c = cand[pos];
                if (c == EMPTY_KEY) {
                    uint64_t old = atomic_cas_u64(&cand[pos], EMPTY_KEY, key);
                    if (old == EMPTY_KEY || old == key) {
                        atomicAdd(&cand_count[pos], 1u);
                        break;
                    }
                    c = old;
                }
                if (c == key) {
                    atomicAdd(&cand_count[pos], 1u);
                    break;
                }
                pos = (pos + 1) & cand_mask;
            }
        }
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

// ------------------------------------------------------------------ phase A: scatter

// Number of `key` in the hot set through the LDS index, or HOT_EMPTY.
__device__ __forceinline__ uint32_t hot_find(const uint32_t* hidx, const uint64_t* hkeys, uint32_t hmask, uint64_t h,
                                             uint64_t key) {
    uint32_t pos = (uint32_t)h & hmask;
    while (true) {
        uint32_t k = hidx[pos];
        if (k == HOT_EMPTY) return HOT_EMPTY;
        if (hkeys[k] == key) return k;
        pos = (pos + 1) & hmask;
    }
}

// LDS bytes of the scatter kernel.
__host__ __device__ inline size_t scatter_lds_bytes(int nhot, int hot_slots, int stride, int nparts) {
    return (size_t)nhot * 8 + (size_t)nhot * stride * 8 + (size_t)hot_slots * 4 + (size_t)nparts * 16;
}

__global__ __launch_bounds__(PA_BLOCK) void part_scatter_kernel(PartArgs a) {
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int S = u.t.stride;
    const int P = a.nparts;
    const int H = a.hot_slots;
    const int K = a.nhot;
    // LDS: hot keys [K], hot states [K][S], hot index [H] (u32), per region: fill, start, cur, first
    uint64_t* hkeys = lds;
    uint64_t* hstate = hkeys + K;
    uint32_t* hidx = (uint32_t*)(hstate + (size_t)K * S);
    uint32_t* fill = hidx + H;   // positions taken in the region's chain of this workgroup
    uint32_t* start = fill + P;  // `fill` at the start of the tile
    uint32_t* cur = start + P;   // the chain's newest block
    uint32_t* first = cur + P;   // first block taken for the tile's records

    for (int k = threadIdx.x; k < K; k += blockDim.x) hkeys[k] = a.hot_keys[k];
    for (int w = threadIdx.x; w < K * S; w += blockDim.x) hstate[w] = u.ident[w % S];
    for (int q = threadIdx.x; q < H; q += blockDim.x) hidx[q] = HOT_EMPTY;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        fill[p] = 0;
        start[p] = 0;
        cur[p] = NO_BLOCK;
        first[p] = NO_BLOCK;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const uint64_t key = hkeys[k];
        if (key == EMPTY_KEY) continue;
        uint32_t pos = (uint32_t)fmix64(key) & (uint32_t)(H - 1);
        while (atomicCAS(&hidx[pos], HOT_EMPTY, (uint32_t)k) != HOT_EMPTY) pos = (pos + 1) & (uint32_t)(H - 1);
    }
    __syncthreads();

    const int RPT = u.rows_per_thread;
    const int64_t T = (int64_t)PA_BLOCK * RPT;
    const int pbits = u.t.pbits;
    const int RW = a.rl.words;
    for (int64_t tile = u.tile_lo + blockIdx.x; tile < u.total_tiles; tile += gridDim.x) {
        const BatchDesc& b = batch_of_tile(u, tile);
        const int64_t row0 = (tile - b.tile_begin) * T;
        uint32_t cold_v[PA_MAX_RPT];  // position in the region's chain
        uint32_t cold_p[PA_MAX_RPT];  // region, NO_BLOCK: the row is not a record
        uint64_t cold_k[PA_MAX_RPT];
#pragma unroll
        for (int r = 0; r < PA_MAX_RPT; r++) {
            cold_p[r] = NO_BLOCK;
            if (r >= RPT) continue;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            bool active = i < b.n;
            if (active && b.sel) active = bit_at(b.sel, i);
            if (active && u.filter_col >= 0) {
                const DCol& f = b.vals[u.filter_col];
                active = is_valid(f, i) && ld_stream((const int32_t*)f.data + i) < u.filter_threshold;
            }
            if (!active) continue;
            uint64_t key, slot;
            const bool keyed = key64_of(u, b, i, &key, &slot);
            if (!keyed || key == EMPTY_KEY) {
                // a reserved row (the null of a 64-bit lane, or a key equal to the empty marker)
                if (keyed) slot = u.t.slots;
                mark_reserved(u.t.counters, keyed ? RES_SENTINEL : RES_NULL);
                apply_ops(u, b, i, u.t.state + slot * (uint64_t)S);
                continue;
            }
            const uint64_t h = fmix64(key);
            if (K) {
                const uint32_t k = hot_find(hidx, hkeys, (uint32_t)(H - 1), h, key);
                if (k != HOT_EMPTY) {
                    apply_ops(u, b, i, hstate + (size_t)k * S);
                    continue;
                }
            }
            const uint32_t p = (uint32_t)region_of_hash(pbits, h);
            cold_p[r] = p;
            cold_k[r] = key;
            cold_v[r] = atomicAdd(&fill[p], 1u);
        }
        __syncthreads();
        // blocks for the positions [start, fill) the tile took in each chain
        for (int p = threadIdx.x; p < P; p += blockDim.x) {
            const uint32_t s0 = start[p], f = fill[p];
            if (f == s0) continue;
            const uint32_t have = (s0 + PART_BLK - 1) / PART_BLK;
            const uint32_t need = (f + PART_BLK - 1) / PART_BLK - have;
            if (need == 0) continue;
            const uint64_t b0 = atomic_add_u64(&u.t.counters[C_POOL], need);
            if (b0 + need > a.max_blocks) {
                atomicOr((unsigned long long*)&u.t.counters[C_ERROR], (unsigned long long)ERR_TABLE_FULL);
                first[p] = NO_BLOCK;
                continue;
            }
            first[p] = (uint32_t)b0;
            for (uint32_t j = 0; j < need; j++) {
                a.block_next[b0 + j] = j ? (uint32_t)(b0 + j - 1) : cur[p];
                a.block_cnt[b0 + j] = PART_BLK;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < PA_MAX_RPT; r++) {
            const uint32_t p = cold_p[r];
            if (p == NO_BLOCK) continue;
            const uint32_t v = cold_v[r];
            const uint32_t have = (start[p] + PART_BLK - 1) / PART_BLK;
            const uint32_t j = v / PART_BLK;
            uint32_t blk;
            if (j < have) {
                blk = cur[p];
            } else {
                blk = first[p];
                if (blk == NO_BLOCK) continue;  // the pool ran out: the launch fails
                blk += j - have;
            }
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            write_record(a.rl, b, i, cold_k[r], a.pool + ((uint64_t)blk * PART_BLK + (v % PART_BLK)) * (uint64_t)RW);
        }
        __syncthreads();
        for (int p = threadIdx.x; p < P; p += blockDim.x) {
            const uint32_t s0 = start[p], f = fill[p];
            if (f == s0) continue;
            const uint32_t have = (s0 + PART_BLK - 1) / PART_BLK;
            const uint32_t need = (f + PART_BLK - 1) / PART_BLK - have;
            if (need && first[p] != NO_BLOCK) cur[p] = first[p] + need - 1;
            start[p] = f;
        }
        __syncthreads();
    }
    // close the chains (the newest block's record count, the heads) and publish the hot partials
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const uint32_t f = fill[p];
        const uint32_t c = cur[p];
        if (f && c != NO_BLOCK) a.block_cnt[c] = f - (f - 1) / PART_BLK * PART_BLK;
        a.heads[(size_t)blockIdx.x * P + p] = f ? c : NO_BLOCK;
    }
    uint64_t* hp = a.hot_part + (size_t)blockIdx.x * K * S;
    for (int w = threadIdx.x; w < K * S; w += blockDim.x) hp[w] = hstate[w];
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

// Entry of `key` in the LDS cache, claiming an empty one; -1 when neither is found within `limit`
// probes (limit 0: no limit -- the caller guarantees a free entry).
__device__ __forceinline__ int cache_find(uint64_t* ckeys, int E, uint64_t h, uint64_t key, int limit) {
    uint32_t pos = (uint32_t)h & (uint32_t)(E - 1);
    for (int q = 0; limit == 0 || q < limit; q++) {
        const uint64_t k = ckeys[pos];
        if (k == key) return (int)pos;
        if (k == EMPTY_KEY) {
            const uint64_t old = atomicCAS((unsigned long long*)&ckeys[pos], (unsigned long long)EMPTY_KEY,
                                           (unsigned long long)key);
            if (old == EMPTY_KEY || old == key) return (int)pos;
        }
        pos = (pos + 1) & (uint32_t)(E - 1);
    }
    return -1;
}

__global__ __launch_bounds__(PB_BLOCK) void part_merge_kernel(PartArgs a) {
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int S = u.t.stride;
    const int E = a.cache_e;
    const int P = a.nparts;
    const uint32_t p = blockIdx.x;
    uint64_t* ckeys = lds;
    uint64_t* cstate = ckeys + E;
    __shared__ int s_flags;  // bit 0: a record found no cache entry, bit 1: a wave has blocks left
    __shared__ int s_stop;
    __shared__ unsigned int s_count, s_inserted, s_kept;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = PB_BLOCK / 64;
    const int RW = a.rl.words;
    uint32_t* rg = u.t.rgroups;
    TableView tv = u.t;
    tv.rgroups = nullptr;  // inserts are counted here, once per merge
    const uint32_t limit = a.region_limit;

    int resume = a.resumed ? a.resume_step[p] : 0;
    if (resume < 0) return;  // merged already

    // Region capacity: can s_count more groups come in?  Counted by every thread before the call;
    // read after the barrier.  Uniform over the workgroup.
    auto room_for_count = [&]() -> bool {
        __syncthreads();
        if (threadIdx.x == 0) s_stop = rg[p] + s_count > limit;
        __syncthreads();
        return !s_stop;
    };
    auto add_inserted = [&]() {  // after a barrier that follows the inserts
        if (threadIdx.x == 0 && s_inserted) {
            rg[p] += s_inserted;
            atomic_add_u64(&u.t.counters[C_GROUPS], s_inserted);
            s_inserted = 0;
        }
    };
    if (threadIdx.x == 0) {
        s_count = 0;
        s_inserted = 0;
    }
    __syncthreads();
    bool stopped = false;
    if (resume == 0) {
        // the hot set's partials of this region, straight into the region (one thread per key)
        for (int k = threadIdx.x; k < a.nhot; k += blockDim.x) {
            const uint64_t key = a.hot_keys[k];
            if (key == EMPTY_KEY || a.hot_total[(size_t)k * S] == 0) continue;
            if (region_of_hash(u.t.pbits, fmix64(key)) == p) atomicAdd(&s_count, 1u);
        }
        if (!room_for_count()) {
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
    int committed = resume;  // steps before committed - 1 are merged (and the hot partials)

    for (int e = threadIdx.x; e < E; e += blockDim.x) ckeys[e] = EMPTY_KEY;
    for (int w = threadIdx.x; w < E * S; w += blockDim.x) cstate[w] = u.ident[w % S];

    // Merges every cache entry that took rows into the region; entries with two rows or more stay
    // cached (at identity) when `keep`.  Returns false, merging nothing, when the region could not
    // take every such entry as a new group.
    auto flush = [&](bool keep) -> bool {
        __syncthreads();
        if (threadIdx.x == 0) {
            s_count = 0;
            s_kept = 0;
        }
        __syncthreads();
        unsigned int mine = 0;
        for (int e = threadIdx.x; e < E; e += blockDim.x)
            if (ckeys[e] != EMPTY_KEY && cstate[(size_t)e * S] != 0) mine++;
        if (mine) atomicAdd(&s_count, mine);
        if (!room_for_count()) return false;
        uint64_t kept[8];
        int nkept = 0;
        for (int e = threadIdx.x; e < E; e += blockDim.x) {
            const uint64_t key = ckeys[e];
            if (key == EMPTY_KEY) continue;
            const uint64_t* l = cstate + (size_t)e * S;
            const uint64_t rows = l[0];
            if (rows) {
                bool ins;
                const uint64_t slot = find_or_insert64(tv, key, &ins);
                if (ins) atomicAdd(&s_inserted, 1u);
                merge_plain(u, l, u.t.state + slot * (uint64_t)S, S);
            }
            if (keep && rows >= 2 && nkept < 8 && atomicAdd(&s_kept, 1u) < (unsigned)(E / 4)) kept[nkept++] = key;
        }
        __syncthreads();
        add_inserted();
        for (int e = threadIdx.x; e < E; e += blockDim.x) ckeys[e] = EMPTY_KEY;
        for (int w = threadIdx.x; w < E * S; w += blockDim.x) cstate[w] = u.ident[w % S];
        __syncthreads();
        for (int q = 0; q < nkept; q++) (void)cache_find(ckeys, E, fmix64(kept[q]), kept[q], 0);
        __syncthreads();
        return true;
    };

    // this wave's chains: wave, wave + nwaves, ... < grid_a
    int chain = wave;
    uint32_t blk = NO_BLOCK;
    auto next_chain = [&]() {
        while (blk == NO_BLOCK && chain < a.grid_a) {
            blk = a.heads[(size_t)chain * P + p];
            chain += nwaves;
        }
    };
    next_chain();
    auto advance = [&]() {
        if (blk == NO_BLOCK) return;
        blk = a.block_next[blk];
        next_chain();
    };
    int step = 0;
    if (!stopped) {
        for (; step < committed - 1; step++) advance();  // merged before the table grew
        __syncthreads();
    }
    while (!stopped) {
        // step `step`: each wave takes one block, lane j its record j
        const uint32_t* rec = nullptr;
        if (blk != NO_BLOCK) {
            if ((uint32_t)lane < a.block_cnt[blk]) rec = a.pool + ((uint64_t)blk * PART_BLK + lane) * (uint64_t)RW;
            advance();
        }
        if (threadIdx.x == 0) s_flags = 0;
        __syncthreads();
        uint64_t key = 0, h = 0;
        int e = -1;
        if (rec) {
            key = *(const uint64_t*)rec;
            h = fmix64(key);
            e = cache_find(ckeys, E, h, key, PB_PROBES);
            if (e < 0) atomicOr(&s_flags, 1);
        }
        if (blk != NO_BLOCK) atomicOr(&s_flags, 2);
        __syncthreads();
        const int flags = s_flags;
        if (flags & 1) {
            // the cache is full: merge it (every step before this one), then place this step's rows
            if (!flush(true)) {
                stopped = true;
                break;
            }
            committed = step + 1;
            if (rec) e = cache_find(ckeys, E, h, key, 0);
        }
        if (rec) apply_src(u, RecSrc{rec, a.rl}, cstate + (size_t)e * S);
        step++;
        if (!(flags & 2)) break;
        __syncthreads();
    }
    if (!stopped) {
        if (flush(false)) committed = -1;
        else stopped = true;
    }
    if (threadIdx.x == 0) {
        a.resume_step[p] = committed;
        if (stopped) atomicOr((unsigned long long*)&u.t.counters[C_PSTOP], 1ull);
    }
}

}  // namespace vr
