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
//                then each of its waves walks its share of the chains on its own (no barrier),
//                folding records into an LDS cache of groups with LDS atomics, and at the end the
//                cache is merged into the region -- new slots claimed by CAS, state rows by plain
//                read-modify-write, since no other workgroup touches the region.  A record whose
//                key finds the cache full around it is marked in a pending map and folded by a
//                further pass after that merge.
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
__device__ const uint32_t k_zero_words[4] = {0, 0, 0, 0};

struct PartArgs {
    UpdateArgs u;
    RecLayout rl;
    ColProg prog[VR_MAX_VALUES];
    int32_t nprog;
    int32_t pad2;
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
    const uint32_t* tile_batch;  // scatter: the batch of every tile
    uint64_t* pending;         // merge: 2 x max_blocks words, the records of a block pending in a pass
    uint64_t max_blocks;
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

// ------------------------------------------------------------------ phase A: scatter

// LDS rows of the hot set and of the merge cache: the key, then the state words, padded to an odd
// number of 8-byte words so that the lanes of a wave updating one word of different rows spread
// over the LDS banks (a 64-byte row would put them all on two).
__host__ __device__ inline int lds_row_words(int stride) { return (stride + 1) | 1; }

// The hot set's LDS index: buckets of HB entries, an entry = 16 bits of the key's hash (bits
// 32-47, not the bucket's) above the key's number in the hot set.  Entries fill each bucket from
// its start, so a bucket with an empty entry ends the search; one bucket read answers almost
// every lookup.
constexpr int HB = 8;
__device__ __forceinline__ uint32_t hot_tag(uint64_t h) { return (uint32_t)(h >> 16) & 0xFFFF0000u; }

// Number of `key` in the hot set, or HOT_EMPTY.
__device__ __forceinline__ uint32_t hot_find(const uint32_t* hidx, const uint64_t* hrow, int HW, uint32_t bmask,
                                             uint64_t h, uint64_t key) {
    const uint32_t tag = hot_tag(h);
    uint32_t b = (uint32_t)h & bmask;
    while (true) {
        const uint4* q = (const uint4*)(hidx + b * HB);
        const uint4 x = q[0], y = q[1];
        const uint32_t e[HB] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
        bool full = true;
#pragma unroll
        for (int j = 0; j < HB; j++) {
            if (e[j] == HOT_EMPTY) full = false;
            else if ((e[j] & 0xFFFF0000u) == tag && hrow[(e[j] & 0xFFFFu) * HW] == key) return e[j] & 0xFFFFu;
        }
        if (!full) return HOT_EMPTY;
        b = (b + 1) & bmask;
    }
}

// LDS bytes of the scatter kernel.
__host__ __device__ inline size_t scatter_lds_bytes(int nhot, int hot_slots, int stride, int nparts) {
    return (size_t)nhot * lds_row_words(stride) * 8 + (size_t)hot_slots * 4 + (size_t)nparts * 16;
}

// Word o of a record held in registers (o uniform): a select per word, no indexed register access.
template <int RW>
__device__ __forceinline__ void put_word(uint32_t (&w)[RW], int o, uint32_t x) {
#pragma unroll
    for (int j = 0; j < RW; j++)
        if (j == o) w[j] = x;
}

// RW: 32-bit words per record (the layout's `words`).  PS: where the column programs come from.
template <int RW, class PS>
__global__ __launch_bounds__(PA_BLOCK) void part_scatter_kernel(PartArgs a) {
    constexpr int MR = PS::max_rpt;  // rows per thread per tile at most (registers per row)
    extern __shared__ uint64_t lds[];
    const UpdateArgs& u = a.u;
    const int S = u.t.stride;
    const int P = a.nparts;
    const int H = a.hot_slots;
    const int K = a.nhot;
    // LDS: hot keys [K], hot states [K][S], hot index [H] (u32), per region: fill, start, cur, first
    const int HW = lds_row_words(S);
    uint64_t* hrow = lds;  // row k: key k of the hot set, its state words
    uint32_t* hidx = (uint32_t*)(hrow + (size_t)K * HW);
    uint32_t* fill = hidx + H;   // positions taken in the region's chain of this workgroup
    uint32_t* start = fill + P;  // `fill` at the start of the tile
    uint32_t* cur = start + P;   // the chain's newest block
    uint32_t* first = cur + P;   // first block taken for the tile's records
    // blocks taken from the pool and not handed out yet: [s_next, s_end); a tile needing more takes
    // a new range [s_alt, s_alt_end) for the regions the rest cannot hold, which then replaces it
    __shared__ uint32_t s_next, s_end, s_need, s_alt, s_alt_end, s_fail;

    for (int w = threadIdx.x; w < K * HW; w += blockDim.x) {
        const int j = w % HW;
        hrow[w] = j == 0 ? a.hot_keys[w / HW] : (j <= S ? u.ident[j - 1] : 0);
    }
    for (int q = threadIdx.x; q < H; q += blockDim.x) hidx[q] = HOT_EMPTY;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        fill[p] = 0;
        start[p] = 0;
        cur[p] = NO_BLOCK;
        first[p] = NO_BLOCK;
    }
    if (threadIdx.x == 0) {
        s_next = 0;
        s_end = 0;
        s_need = 0;
        s_alt = 0;
        s_alt_end = 0;
        s_fail = 0;
    }
    __syncthreads();
    const uint32_t hbmask = (uint32_t)(H / HB - 1);
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
    __syncthreads();

    const int RPT = u.rows_per_thread;
    const int64_t T = (int64_t)PA_BLOCK * RPT;
    const int pbits = u.t.pbits;
    // Tiles are grid-strided, so that the CUs read neighbouring addresses at any time.  The batch
    // descriptor of a tile is in LDS while the tile runs (its column pointers are read there once
    // per column and tile); the next tile's is fetched meanwhile into the other buffer.
    constexpr int DESC_WORDS = (int)(sizeof(BatchDesc) / 8);
    __shared__ BatchDesc lb[2];
    int64_t tile = u.tile_lo + blockIdx.x;
    if (tile < u.total_tiles) {
        const uint64_t* src = (const uint64_t*)&batch_at(u.batches, a.tile_batch[tile]);
        for (int w = threadIdx.x; w < DESC_WORDS; w += blockDim.x) ((uint64_t*)&lb[0])[w] = src[w];
    }
    __syncthreads();
    int cur_buf = 0;
    for (; tile < u.total_tiles; tile += gridDim.x) {
        const int64_t ntile = tile + gridDim.x;
        const bool more = ntile < u.total_tiles;
        uint64_t next_word = 0;
        if (more && threadIdx.x < DESC_WORDS)
            next_word = ((const uint64_t*)&batch_at(u.batches, a.tile_batch[ntile]))[threadIdx.x];
        const BatchDesc& b = lb[cur_buf];
        const int64_t row0 = (tile - b.tile_begin) * T;
        // per row: key, and its fate -- ent >= 0: hot entry; -1: nothing (inactive or folded);
        // -2: a record of region pv at chain position vv
        uint64_t kk[MR];
        int32_t ent[MR];
        uint32_t pv[MR], vv[MR];
        bool nul[MR];
        // Every load of the tile first, with no branch between them (one round trip per tile): the
        // row index is clamped into the batch, a lane is read as 32-bit words at i * width (the
        // second word re-reads the first for a 4-byte lane), bitmaps the batch lacks are read from
        // k_all_ones.  Validity, selection and the filter are applied afterwards.
        uint32_t vm[MR];
        uint64_t xa[MR][PART_MAX_PROG], xb[MR][PART_MAX_PROG];
        uint32_t vbyte[MR][PART_MAX_PROG];
        uint32_t lo[MR][PART_MAX_PROG], hi[MR][PART_MAX_PROG];
        uint32_t w2[MR][PART_MAX_PROG], w3[MR][PART_MAX_PROG];
        uint32_t sbyte[MR], fbyte[MR], kbyte[MR];
        int32_t fval[MR];
        uint32_t klo[MR], khi[MR];
        const bool fast_key = u.ks.mode == KM_I64;
        const bool has_filter = u.filter_col >= 0;
        const DCol& fcol = b.vals[has_filter ? u.filter_col : 0];
        const DCol& kcol = b.keys[0];
#pragma unroll
        for (int r = 0; r < MR; r++) {
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            const int64_t ci = i < b.n ? i : b.n - 1;
            sbyte[r] = *gptr(b.sel ? b.sel + (ci >> 3) : k_all_ones);
            fbyte[r] = *gptr(has_filter && fcol.valid ? fcol.valid + (ci >> 3) : k_all_ones);
            fval[r] = (int32_t)*gptr(has_filter ? (const uint32_t*)fcol.data + ci : k_zero_words);
            if (fast_key) {
                kbyte[r] = *gptr(kcol.valid ? kcol.valid + (ci >> 3) : k_all_ones);
                const uint32_t* kp = (const uint32_t*)kcol.data + 2 * ci;
                klo[r] = *gptr(kp);
                khi[r] = *gptr(kp + 1);
            }
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++) {
                const bool on = q < PS::nprog(a);
                const ColProg& cp = PS::prog(a, on ? q : 0);
                const DCol& d = b.vals[cp.col];
                vbyte[r][q] = *gptr(on && d.valid ? d.valid + (ci >> 3) : k_all_ones);
                const int wd = on ? cp.width : 0;
                const uint32_t* vp = wd ? (const uint32_t*)((const uint8_t*)d.data + ci * wd) : k_zero_words;
                lo[r][q] = *gptr(vp);
                hi[r][q] = *gptr(vp + (wd >= 8 ? 1 : 0));
                w2[r][q] = *gptr(vp + (wd >= 16 ? 2 : 0));
                w3[r][q] = *gptr(vp + (wd >= 16 ? 3 : 0));
            }
        }
#pragma unroll
        for (int r = 0; r < MR; r++) {
            ent[r] = -1;
            kk[r] = 0;
            nul[r] = false;
            pv[r] = 0;
            vv[r] = 0;
            vm[r] = 0;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            const int bit = (int)(i & 7);
#pragma unroll
            for (int q = 0; q < PART_MAX_PROG; q++) {
                const int wd = q < PS::nprog(a) ? PS::prog(a, q).width : 0;
                xa[r][q] = (uint64_t)lo[r][q] | (wd >= 8 ? (uint64_t)hi[r][q] << 32 : 0ull);
                xb[r][q] = wd >= 16 ? ((uint64_t)w2[r][q] | ((uint64_t)w3[r][q] << 32)) : 0ull;
                if (q < PS::nprog(a) && ((vbyte[r][q] >> bit) & 1u)) vm[r] |= 1u << q;
            }
            if (r >= RPT || i >= b.n || !((sbyte[r] >> bit) & 1u)) continue;
            if (has_filter && !(((fbyte[r] >> bit) & 1u) && fval[r] < u.filter_threshold)) continue;
            if (fast_key) {
                nul[r] = !((kbyte[r] >> bit) & 1u);
                kk[r] = (uint64_t)klo[r] | ((uint64_t)khi[r] << 32);
            } else {
                uint64_t slot;
                nul[r] = !key64_of(u, b, i, &kk[r], &slot);
            }
            ent[r] = -3;  // active, not classified yet
        }
#pragma unroll
        for (int r = 0; r < MR; r++) {
            if (ent[r] != -3) continue;
            const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
            const uint64_t key = kk[r];
            if (nul[r] || key == EMPTY_KEY) {
                // a reserved row (the null of a 64-bit lane, or a key equal to the empty marker)
                const uint64_t slot = nul[r] ? u.t.slots + 1 : u.t.slots;
                mark_reserved(u.t.counters, nul[r] ? RES_NULL : RES_SENTINEL);
                apply_ops(u, b, i, u.t.state + slot * (uint64_t)S);
                ent[r] = -1;
                continue;
            }
            const uint64_t h = fmix64(key);
            if (K) {
                const uint32_t k = hot_find(hidx, hrow, HW, hbmask, h, key);
                if (k != HOT_EMPTY) {
                    ent[r] = (int32_t)k;
                    continue;
                }
            }
            const uint32_t p = (uint32_t)region_of_hash(pbits, h);
            pv[r] = p;
            vv[r] = atomicAdd(&fill[p], 1u);
            ent[r] = -2;
        }
        __syncthreads();
        // blocks for the positions [start, fill) the tile took in each chain, from the
        // workgroup's share of the pool (refilled with one device atomic per PA_CHUNK blocks; the
        // one region whose blocks straddle the end of the old share leaves under PA_GAP unused)
        for (int p = threadIdx.x; p < P; p += blockDim.x) {
            const uint32_t need = (fill[p] + PART_BLK - 1) / PART_BLK - (start[p] + PART_BLK - 1) / PART_BLK;
            if (need) atomicAdd(&s_need, need);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t need = s_need;
            if (!s_fail && s_end - s_next < need) {
                const uint32_t take = need > (uint32_t)PA_CHUNK ? need : (uint32_t)PA_CHUNK;
                const uint64_t b0 = atomic_add_u64(&u.t.counters[C_POOL], take);
                if (b0 + take > a.max_blocks) {
                    atomicOr((unsigned long long*)&u.t.counters[C_ERROR], (unsigned long long)ERR_TABLE_FULL);
                    s_fail = 1;  // nothing handed out from here on: the launch fails
                } else {
                    s_alt = (uint32_t)b0;
                    s_alt_end = (uint32_t)(b0 + take);
                }
            }
            s_need = 0;
        }
        __syncthreads();
        const bool refilled = s_alt_end != 0;  // else the share holds the whole tile
        for (int p = threadIdx.x; p < P; p += blockDim.x) {
            const uint32_t have = (start[p] + PART_BLK - 1) / PART_BLK;
            const uint32_t need = (fill[p] + PART_BLK - 1) / PART_BLK - have;
            if (need == 0) continue;
            if (s_fail) {
                first[p] = NO_BLOCK;
                continue;
            }
            uint32_t b0 = atomicAdd(&s_next, need);
            // the rest of the old share cannot hold it: the new range (which holds the whole tile)
            if (refilled && b0 + need > s_end) b0 = atomicAdd(&s_alt, need);
            first[p] = b0;
            for (uint32_t j = 0; j < need; j++) {
                a.block_next[b0 + j] = j ? b0 + j - 1 : cur[p];
                a.block_cnt[b0 + j] = PART_BLK;
            }
        }
        __syncthreads();
        // record addresses of the cold rows
        uint32_t* rec[MR];
        uint32_t vbits[MR];
#pragma unroll
        for (int r = 0; r < MR; r++) {
            rec[r] = nullptr;
            vbits[r] = 0;
            if (ent[r] != -2) continue;
            const uint32_t p = pv[r], v = vv[r];
            const uint32_t have = (start[p] + PART_BLK - 1) / PART_BLK;
            const uint32_t j = v / PART_BLK;
            uint32_t blk = j < have ? cur[p] : first[p];
            if (blk == NO_BLOCK) {  // the pool ran out: the launch fails
                ent[r] = -1;
                continue;
            }
            if (j >= have) blk += j - have;
            rec[r] = a.pool + ((uint64_t)blk * PART_BLK + (v % PART_BLK)) * (uint64_t)RW;
        }
        // COUNT(*): word 0 of every state row
#pragma unroll
        for (int r = 0; r < MR; r++)
            if (ent[r] >= 0) atomic_add_u64(hrow + (size_t)ent[r] * HW + 1, 1);
        // each column's program on the hot rows, the column's field of the records
#pragma unroll
        for (int q = 0; q < PART_MAX_PROG; q++) {
            if (q >= PS::nprog(a)) break;
            const ColProg cp = PS::prog(a, q);
            const int o = cp.rec_off;
#pragma unroll
            for (int r = 0; r < MR; r++) {
                const bool v = (vm[r] >> q) & 1u;
                if (ent[r] >= 0) {
                    uint64_t* e = hrow + (size_t)ent[r] * HW + 1;
                    if (cp.flags & CP_GENERIC) {
                        const int64_t i = row0 + (int64_t)r * PA_BLOCK + threadIdx.x;
                        const RegSrc src{v, xa[r][q], xb[r][q], (uint64_t)(b.row_base + i), cp.type};
                        for (int k = 0; k < u.nops; k++)
                            if (u.ops[k].col == cp.col) apply_one(u, u.ops[k], src, e);
                    } else {
                        apply_prog(u, cp, v, xa[r][q], xb[r][q], e);
                    }
                } else if (ent[r] == -2) {
                    if (v) vbits[r] |= 1u << cp.col;
                    if (o < 0 || !v) continue;
                    uint32_t* dst = rec[r] + o;
                    if (cp.type == VR_INT32) {
                        *dst = (uint32_t)xa[r][q];
                    } else {
                        *(uint64_t*)dst = xa[r][q];
                        if (cp.type == VR_DECIMAL128) *(uint64_t*)(dst + 2) = xb[r][q];
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < MR; r++) {
            if (ent[r] != -2) continue;
            *(uint64_t*)rec[r] = kk[r];
            rec[r][PS::valid_off(a)] = vbits[r];
            if (PS::row_off(a) >= 0)
                *(uint64_t*)(rec[r] + PS::row_off(a)) = (uint64_t)(b.row_base + row0 + (int64_t)r * PA_BLOCK + threadIdx.x);
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
        if (more && threadIdx.x < DESC_WORDS) ((uint64_t*)&lb[cur_buf ^ 1])[threadIdx.x] = next_word;
        if (threadIdx.x == 0 && s_alt_end) {  // the new range is the share from here on
            s_next = s_alt;
            s_end = s_alt_end;
            s_alt = s_alt_end = 0;
        }
        __syncthreads();
        cur_buf ^= 1;
    }
    // close the chains (the newest block's record count, the heads) and publish the hot partials
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const uint32_t f = fill[p];
        const uint32_t c = cur[p];
        if (f && c != NO_BLOCK) a.block_cnt[c] = f - (f - 1) / PART_BLK * PART_BLK;
        a.heads[(size_t)blockIdx.x * P + p] = f ? c : NO_BLOCK;
    }
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

// LDS bytes of the merge: the cache (E keys and state rows) and the waves' block lists.
__host__ __device__ inline size_t merge_lds_bytes(int E, int stride) {
    return (size_t)E * lds_row_words(stride) * 8 + (size_t)PB_GLIST * 4;
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
    const int P = a.nparts;
    const uint32_t p = blockIdx.x;
    const int CW = lds_row_words(S);
    uint64_t* crow = lds;  // cache entry e: row e (key, state words)
    const int lane = threadIdx.x & 63;
    // the workgroup's list of blocks: (records - 1) << PB_ID_BITS | block
    uint32_t* glist = (uint32_t*)(crow + (size_t)E * CW);
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

    // The region's blocks, handed out to the waves as they ask (the waves' work evens out whatever
    // the chains' lengths).  A round: the threads owning a chain (thread c: chain c < grid_a) walk
    // it, at most `cap` blocks each, into the workgroup's block list; after a barrier the waves
    // take blocks from the list with an LDS counter; rounds repeat until every chain is walked.
    __shared__ unsigned int s_gcount, s_gnext, s_walk;
    const int cap = PB_GLIST / a.grid_a < PB_LANE_BLOCKS ? PB_GLIST / a.grid_a : PB_LANE_BLOCKS;
    uint32_t ccur = NO_BLOCK;
    auto walk = [&]() {
        for (int t = 0; t < cap && ccur != NO_BLOCK; t++) {
            const uint32_t b = ccur;
            const uint32_t n = a.block_cnt[b];
            ccur = a.block_next[b];
            glist[atomicAdd(&s_gcount, 1u)] = ((n - 1) << PB_ID_BITS) | b;
        }
        if (ccur != NO_BLOCK) s_walk = 1;
    };
    // the next block of the list for this wave (NO_BLOCK: none) and its record count
    auto fetch = [&](uint32_t* n) -> uint32_t {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&s_gnext, 1u);
        i = __shfl(i, 0, 64);
        if (i >= s_gcount) {
            *n = 0;
            return NO_BLOCK;
        }
        const uint32_t x = glist[i];
        *n = (x >> PB_ID_BITS) + 1;
        return x & ((1u << PB_ID_BITS) - 1);
    };
    if (threadIdx.x == 0) {
        s_gcount = 0;
        s_gnext = 0;
        s_walk = 0;
    }
    __syncthreads();

    // Passes.  Pass 0 folds every record; a record that finds no cache entry (the cache is full)
    // is marked in its block's word of the other pending map, and pass k >= 1 folds the records
    // marked by pass k - 1.  Each pass ends with one merge of the cache into the region.  The two
    // maps are used in turn, so a pass whose merge finds the region short stops with its input
    // intact and resumes there after the table grows.  Within a round the waves run on their own
    // (no barrier): each loads the records of its blocks two blocks ahead, into three buffers
    // used in turn (each unrolled step names its buffers, so no register copy waits for a load in
    // flight).
    int pass = resume - 1;
    uint32_t ra[RW], rb[RW], rc[RW];
    while (!stopped) {
        const uint64_t* rd = pass ? a.pending + (size_t)(pass & 1) * a.max_blocks : nullptr;
        uint64_t* wr = a.pending + (size_t)((pass + 1) & 1) * a.max_blocks;
        ccur = threadIdx.x < a.grid_a ? a.heads[(size_t)threadIdx.x * P + p] : NO_BLOCK;
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
        while (true) {  // rounds
            walk();
            __syncthreads();
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
            __syncthreads();
            const bool again = s_walk != 0;
            __syncthreads();
            if (threadIdx.x == 0) {
                s_gcount = 0;
                s_gnext = 0;
                s_walk = 0;
            }
            __syncthreads();
            if (!again) break;
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
