// Group-by hash aggregation on gfx950: the vr_agg_* entry points of include/vr.h.
//
// Replaces, per batch, GroupKeyTable.assign + GroupedAccumulators.*.update (reference:
// kernels/src/main/java/io/vecruntime/kernels/GroupKeyTable.java:709, GroupedAccumulators.java:224,
// 380,652,1017,1212,1287) as one fused launch, and the emission GroupKeyTable.writeKeys +
// GroupedAggState.writeBuffer (GroupKeyTable.java:2851, spark/.../agg/VectorAggregates.scala:63).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <future>
#include <memory>

#include "vr_agg_part.h"
#include "vr_internal.h"
#include "vr_jit.h"
#include "vr_strings.h"

namespace vr {

// ------------------------------------------------------------------ kernels

constexpr int AGG_BLOCK = 256;
constexpr int MAX_KEEP = 8;  // LDS entries one thread re-inserts after an eviction (E <= 2048)

// Index of `key` in the LDS cache (claiming an empty entry when `insert`), -1 when absent or full.
__device__ __forceinline__ int lds_find(uint64_t* lkeys, int E, uint64_t key, bool insert) {
    uint32_t h = (uint32_t)(fmix64(key)) & (uint32_t)(E - 1);
    for (int p = 0; p < 16; p++) {
        uint64_t k = lkeys[h];
        if (k == key) return (int)h;
        if (k == EMPTY_KEY) {
            if (!insert) return -1;
            uint64_t old = atomicCAS((unsigned long long*)&lkeys[h], (unsigned long long)EMPTY_KEY,
                                     (unsigned long long)key);
            if (old == EMPTY_KEY || old == key) return (int)h;
        }
        h = (h + 1) & (uint32_t)(E - 1);
    }
    return -1;
}

__device__ __forceinline__ void lds_clear_state(const UpdateArgs& a, uint64_t* lstate, int E, int S) {
    for (int w = threadIdx.x; w < E * S; w += blockDim.x) lstate[w] = a.ident[w % S];
}

// Flushes every LDS entry into its global row.  With `evict`, entries that took fewer than two
// rows since the last flush give their place up; the others stay cached (with cleared partials).
// `by_key`: entries are keyed by the group's key (the global slot is found here), else by slot.
__device__ __forceinline__ void lds_flush(const UpdateArgs& a, uint64_t* lkeys, uint64_t* lstate, int E, bool by_key,
                                          bool evict, unsigned int* inserted) {
    const int S = a.t.stride;
    uint64_t keep[MAX_KEEP];
    int nkeep = 0;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        uint64_t k = lkeys[e];
        if (k == EMPTY_KEY) continue;
        uint64_t* l = lstate + (size_t)e * S;
        uint64_t slot = k;
        if (by_key) {
            bool ins;
            slot = find_or_insert64(a.t, k, &ins);
            if (ins) atomicAdd(inserted, 1u);
        }
        bool hot = l[0] >= 2;  // word 0 is the group's ROWS count
        if (l[0]) flush_entry(a, l, a.t.state + slot * (uint64_t)S);
        if (evict && hot && nkeep < MAX_KEEP) keep[nkeep++] = k;
    }
    if (!evict) return;
    __syncthreads();
    for (int e = threadIdx.x; e < E; e += blockDim.x) lkeys[e] = EMPTY_KEY;
    lds_clear_state(a, lstate, E, S);
    __syncthreads();
    for (int q = 0; q < nkeep; q++) (void)lds_find(lkeys, E, keep[q], true);
    __syncthreads();
}

// Locates the batch of a tile in a batched launch (tile_begin ascending).
__device__ __forceinline__ int find_batch(const BatchDesc* b, int nb, int64_t tile) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (batch_at(b, mid).tile_begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool MULTI>
__global__ __launch_bounds__(AGG_BLOCK) void agg_update_kernel(UpdateArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ unsigned int inserted;
    const int E = a.lds_entries;
    const int S = a.t.stride;
    uint64_t* lkeys = lds;
    uint64_t* lstate = lds + E;
    // 64-bit keys are cached by key (an LDS hit touches no global memory); wide keys, and every
    // key of an update that records each row's slot (for an ordered fold), by slot.
    const bool by_key = a.ks.mode != KM_WIDE && !a.det && !a.row_slot;
    if (threadIdx.x == 0) inserted = 0;
    if (E) {
        for (int e = threadIdx.x; e < E; e += blockDim.x) lkeys[e] = EMPTY_KEY;
        lds_clear_state(a, lstate, E, S);
    }
    __syncthreads();
    const int64_t T = (int64_t)blockDim.x * a.rows_per_thread;
    const bool guard = a.resume != nullptr;
    __shared__ int stop;
    int since_flush = 0;
    int64_t tile = a.resumed ? a.resume[blockIdx.x] : a.tile_lo + blockIdx.x;
    for (; tile < a.total_tiles; tile += gridDim.x) {
        if (guard) {
            __syncthreads();  // the previous tile's inserts are all in `inserted`
            if (threadIdx.x == 0) {
                if (inserted) atomic_add_u64(&a.t.counters[C_GROUPS], inserted);
                inserted = 0;
                stop = ld_coherent(&a.t.counters[C_GROUPS]) >= a.soft_limit;
            }
            __syncthreads();
            if (stop) break;
        }
        const BatchDesc* bp;
        int64_t row0;
        if (MULTI) {
            bp = &batch_at(a.batches, find_batch(a.batches, a.num_batches, tile));
            row0 = (tile - bp->tile_begin) * T;
        } else {
            bp = &a.one;
            row0 = tile * T;
        }
        const BatchDesc& b = *bp;
        const int64_t n = b.n;
        for (int r = 0; r < a.rows_per_thread; r++) {
            int64_t i = row0 + (int64_t)r * blockDim.x + threadIdx.x;
            bool active = i < n;
            if (active && b.sel) active = bit_at(b.sel, i);
            if (active && a.filter_col >= 0) {
                const DCol& f = b.vals[a.filter_col];
                active = is_valid(f, i) && ld_stream((const int32_t*)f.data + i) < a.filter_threshold;
            }
            uint64_t slot = ~0ull;
            if (active) {
                bool ins = false;
                uint64_t* row = nullptr;
                if (by_key) {
                    uint64_t key;
                    bool keyed = key64_of(a, b, i, &key, &slot);
                    if (keyed && key != EMPTY_KEY && E) {
                        int e = lds_find(lkeys, E, key, true);
                        if (e >= 0) row = lstate + (size_t)e * S;
                    }
                    if (!row) {
                        if (keyed) slot = find_or_insert64(a.t, key, &ins);
                        else mark_reserved(a.t.counters, RES_NULL);
                        row = a.t.state + slot * (uint64_t)S;
                    }
                } else {
                    slot = row_slot_of(a, b, i, &ins, true);
                    if (E) {
                        int e = lds_find(lkeys, E, slot, true);
                        if (e >= 0) row = lstate + (size_t)e * S;
                    }
                    if (!row) row = a.t.state + slot * (uint64_t)S;
                }
                if (ins) atomicAdd(&inserted, 1u);
                apply_ops(a, b, i, row);
            }
            if (!MULTI && a.row_slot && i < n) a.row_slot[i] = active ? (uint32_t)slot : 0xFFFFFFFFu;
        }
        if (E && ++since_flush >= a.flush_tiles) {
            __syncthreads();
            lds_flush(a, lkeys, lstate, E, by_key, true, &inserted);
            since_flush = 0;
        }
    }
    __syncthreads();
    if (E && since_flush) lds_flush(a, lkeys, lstate, E, by_key, false, &inserted);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (inserted) atomic_add_u64(&a.t.counters[C_GROUPS], inserted);
        if (guard) {
            a.resume[blockIdx.x] = tile;
            if (tile < a.total_tiles) atomicOr((unsigned long long*)&a.t.counters[C_STOP], 1ull);
        }
    }
}

// Fills state rows [0, rows) with the identity row.  A wave covers 64 consecutive rows and its
// lanes store consecutive words, so the stores coalesce whatever the row width.
__global__ void init_state_kernel(uint64_t* state, uint64_t rows, UpdateArgs a) {
    const int S = a.t.stride;
    const int lane = threadIdx.x & 63;
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t wave = (((uint64_t)blockIdx.x * blockDim.x) + threadIdx.x) >> 6;
    for (uint64_t r0 = wave * 64; r0 < rows; r0 += waves * 64) {
        const uint64_t nrows = rows - r0 < 64 ? rows - r0 : 64;
        uint64_t* d = state + r0 * (uint64_t)S;
        const int words = (int)nrows * S;
        for (int j = lane; j < words; j += 64) d[j] = a.ident[j % S];
    }
}

// Resets the listed rows (an emission order): key cells empty, state at identity.
__global__ void reset_rows_kernel(const uint64_t* order, int64_t m, uint64_t* cells, int W, uint64_t slots, uint64_t empty,
                                  UpdateArgs a) {
    const int S = a.t.stride;
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < m; o += (int64_t)gridDim.x * blockDim.x) {
        uint64_t r = order[o];
        if (r < slots)
            for (int k = 0; k < W; k++) cells[r * (uint64_t)W + k] = empty;
        uint64_t* d = a.t.state + r * (uint64_t)S;
        for (int k = 0; k < S; k++) d[k] = a.ident[k];
    }
}

// Inserts `key` into a table of 64-bit keys (its state row is at identity already).
__global__ void insert_key_kernel(TableView t, uint64_t key) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    bool ins;
    (void)find_or_insert64(t, key, &ins);
    if (ins) atomic_add_u64(&t.counters[C_GROUPS], 1);
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Moves every group of `from` into the (larger, empty) table of `a`.
__global__ void migrate_kernel(TableView from, UpdateArgs a) {
    const TableView& to = a.t;
    const int W = from.width;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < from.slots + 2;
         s += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t dst;
        if (s >= from.slots) {
            uint64_t bit = (s == from.slots) ? RES_SENTINEL : RES_NULL;
            if (!(from.counters[C_RESERVED] & bit)) continue;
            dst = to.slots + (s - from.slots);
        } else if (a.ks.mode != KM_WIDE) {
            uint64_t key = from.cells[s];
            if (key == EMPTY_KEY) continue;
            bool ins;
            dst = find_or_insert64(to, key, &ins);
        } else {
            const uint64_t* c = from.cells + s * (uint64_t)W;
            if (c[0] == 0) continue;
            uint32_t w[4 * VR_MAX_KEYS + 1];
            for (int i = 0; i < W; i++) w[i] = (uint32_t)c[i];
            bool ins;
            dst = find_or_insert_wide(to, w, W, &ins, true);
        }
        const uint64_t* src = from.state + s * (uint64_t)from.stride;
        uint64_t* d = to.state + dst * (uint64_t)to.stride;
        for (int k = 0; k < to.stride; k++) d[k] = src[k];
    }
}

// After a radix sort of (slot, row) pairs by slot, the first position of each slot's run folds that
// run's values in row order onto the group's state:
//  * deterministic mode, double sums: Spark's `sum += x` chain across batches
//    (GroupedAccumulators.DoubleSum(strict), GroupedAccumulators.java:202-206,276-280);
//  * 128-bit MIN/MAX (state lo, hi, any): Decimal128MinMax.offer -- the first value, then one
//    strictly better by signed 128-bit order (GroupedAccumulators.java:1410-1422, Decimal128.java:145).
__global__ void ordered_fold_kernel(const uint32_t* slots, const uint32_t* rows, int64_t n, UpdateArgs a) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        uint32_t s = slots[p];
        if (s == 0xFFFFFFFFu || (p > 0 && slots[p - 1] == s)) continue;
        uint64_t* g = a.t.state + (uint64_t)s * a.t.stride;
        for (int k = 0; k < a.nops; k++) {
            const Op& op = a.ops[k];
            const DCol& c = a.one.vals[op.col];
            if (op.code == OP_SUM_F64 && a.det) {
                double acc = __longlong_as_double((long long)g[op.off]);
                for (int64_t q = p; q < n && slots[q] == s; q++) {
                    uint32_t i = rows[q];
                    if (is_valid(c, i)) acc += ((const double*)c.data)[i];
                }
                g[op.off] = (uint64_t)__double_as_longlong(acc);
            } else if (op.code == OP_SUM_I64 && ((a.checked_sums >> k) & 1u)) {
                // LongSum(checked) (GroupedAccumulators.java:563-601): Math.addExact per row, so any
                // prefix of the group's rows in row order outside the long range is an error.  The
                // state already holds this batch's rows: the total before them is the state's exact
                // value less their sum.
                const bool i32 = c.type == VR_INT32;
                __int128 run = 0;
                for (int64_t q = p; q < n && slots[q] == s; q++) {
                    const uint32_t i = rows[q];
                    if (is_valid(c, i)) run += i32 ? (int64_t)((const int32_t*)c.data)[i] : ((const int64_t*)c.data)[i];
                }
                __int128 acc = (__int128)g[op.off] + ((__int128)(int64_t)g[op.off + 1] << 32) - run;
                bool over = acc < (__int128)INT64_MIN || acc > (__int128)INT64_MAX;
                for (int64_t q = p; q < n && slots[q] == s && !over; q++) {
                    const uint32_t i = rows[q];
                    if (!is_valid(c, i)) continue;
                    acc += i32 ? (int64_t)((const int32_t*)c.data)[i] : ((const int64_t*)c.data)[i];
                    over = acc < (__int128)INT64_MIN || acc > (__int128)INT64_MAX;
                }
                if (over) atomicOr((unsigned long long*)&a.t.counters[C_ERROR], (unsigned long long)ERR_LONG_OVERFLOW);
            } else if (op.code == OP_MIN_D128 || op.code == OP_MAX_D128) {
                const bool mn = op.code == OP_MIN_D128;
                uint64_t lo = g[op.off], hi = g[op.off + 1], any = g[op.off + 2];
                for (int64_t q = p; q < n && slots[q] == s; q++) {
                    const uint32_t i = rows[q];
                    if (!is_valid(c, i)) continue;
                    const uint64_t xl = ((const uint64_t*)c.data)[2 * (uint64_t)i];
                    const uint64_t xh = ((const uint64_t*)c.data)[2 * (uint64_t)i + 1];
                    const int cmp = xh != hi ? ((int64_t)xh < (int64_t)hi ? -1 : 1) : (xl == lo ? 0 : (xl < lo ? -1 : 1));
                    if (!any || (mn ? cmp < 0 : cmp > 0)) {
                        lo = xl;
                        hi = xh;
                        any = 1;
                    }
                }
                g[op.off] = lo;
                g[op.off + 1] = hi;
                g[op.off + 2] = any;
            }
        }
    }
}

__global__ void iota_u32_kernel(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// Records the payload of each group's first NaN (the NaN a tie keeps), for the updates that saw one.
__device__ __forceinline__ void nan_payload_rows(const UpdateArgs& a, const BatchDesc& b, int64_t i0, int64_t i1,
                                                 int64_t step) {
    for (int64_t i = i0; i < i1; i += step) {
        if (b.sel && !bit_at(b.sel, i)) continue;
        if (a.filter_col >= 0) {
            const DCol& f = b.vals[a.filter_col];
            if (!is_valid(f, i) || ((const int32_t*)f.data)[i] >= a.filter_threshold) continue;
        }
        uint64_t slot = ~0ull;
        for (int k = 0; k < a.nops; k++) {
            const Op& op = a.ops[k];
            if (op.code != OP_MIN_F64 && op.code != OP_MAX_F64) continue;
            const DCol& c = b.vals[op.col];
            if (!is_valid(c, i)) continue;
            uint64_t bits = ((const uint64_t*)c.data)[i];
            if (!is_nan_bits(bits)) continue;
            if (slot == ~0ull) {
                bool ins;
                slot = row_slot_of(a, b, i, &ins, false);
                // a key the table does not hold went to the spare row when its region was full
                // (the update then fails with the table-full error)
                if (slot == ~0ull) slot = a.t.slots + 2;
            }
            uint64_t* w = a.t.state + slot * (uint64_t)a.t.stride + op.off;
            if (w[2] == (uint64_t)(b.row_base + i)) w[3] = bits;
        }
    }
}

template <bool MULTI>
__global__ void nan_payload_kernel(UpdateArgs a) {
    if (a.t.counters[C_NAN] != (uint64_t)a.epoch) return;
    if (!MULTI) {  // the rows of tiles [tile_lo, total_tiles) of the update launch
        const int64_t T = (int64_t)AGG_BLOCK * a.rows_per_thread;
        const int64_t hi = a.total_tiles * T < a.one.n ? a.total_tiles * T : a.one.n;
        nan_payload_rows(a, a.one, a.tile_lo * T + (int64_t)blockIdx.x * blockDim.x + threadIdx.x, hi,
                         (int64_t)gridDim.x * blockDim.x);
        return;
    }
    const int64_t T = (int64_t)blockDim.x * a.rows_per_thread;
    for (int64_t tile = a.tile_lo + blockIdx.x; tile < a.total_tiles; tile += gridDim.x) {
        const BatchDesc& b = batch_at(a.batches, find_batch(a.batches, a.num_batches, tile));
        int64_t row0 = (tile - b.tile_begin) * T;
        int64_t row1 = row0 + T < b.n ? row0 + T : b.n;
        nan_payload_rows(a, b, row0 + threadIdx.x, row1, blockDim.x);
    }
}

// ---- emission order: occupied slots in slot order, by a count / scan / write compaction

constexpr int ORDER_CHUNK = 4096;  // slots per block

__device__ __forceinline__ bool occupied(const uint64_t* cells, int W, uint64_t s) {
    uint64_t c = cells[s * (uint64_t)W];
    return W == 1 ? c != EMPTY_KEY : c != 0;
}

__global__ __launch_bounds__(256) void count_occupied_kernel(const uint64_t* cells, int W, uint64_t slots,
                                                             uint64_t* block_counts) {
    typedef hipcub::BlockReduce<uint32_t, 256> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    uint64_t base = (uint64_t)blockIdx.x * ORDER_CHUNK;
    uint32_t c = 0;
    for (int k = threadIdx.x; k < ORDER_CHUNK; k += 256) {
        uint64_t s = base + k;
        if (s < slots && occupied(cells, W, s)) c++;
    }
    uint32_t total = Reduce(tmp).Sum(c);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void write_occupied_kernel(const uint64_t* cells, int W, uint64_t slots,
                                                             const uint64_t* block_offsets, uint64_t* order,
                                                             uint64_t cap) {
    typedef hipcub::BlockScan<uint32_t, 256> Scan;
    __shared__ typename Scan::TempStorage tmp;
    uint64_t base = (uint64_t)blockIdx.x * ORDER_CHUNK;
    uint64_t out = block_offsets[blockIdx.x];
    for (int k0 = 0; k0 < ORDER_CHUNK; k0 += 256) {  // slot order: a 256-slot window at a time
        uint64_t s = base + k0 + threadIdx.x;
        uint32_t flag = (s < slots && occupied(cells, W, s)) ? 1u : 0u;
        uint32_t pos, total;
        Scan(tmp).ExclusiveSum(flag, pos, total);
        if (flag && out + pos < cap) order[out + pos] = s;
        out += total;
        __syncthreads();
    }
}

// Checks the order's length (the scan's total) against the group count and appends the reserved
// rows (the sentinel key's, the null key's) after the occupied slots.
__global__ void order_tail_kernel(const uint64_t* block_counts, const uint64_t* block_offsets, uint64_t nblocks,
                                  uint64_t selected, uint64_t* order, uint64_t x0, uint64_t x1, int nx,
                                  uint64_t* counters) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (block_counts[nblocks - 1] + block_offsets[nblocks - 1] != selected)
        atomicOr((unsigned long long*)&counters[C_ERROR], (unsigned long long)ERR_ORDER);
    if (nx > 0) order[selected] = x0;
    if (nx > 1) order[selected + 1] = x1;
}

__global__ void first_rows_kernel(const uint64_t* order, int64_t m, const uint64_t* state, int stride, int off,
                                  uint64_t* out) {
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < m; o += (int64_t)gridDim.x * blockDim.x)
        out[o] = state[order[o] * (uint64_t)stride + off];
}

// ------------------------------------------------------------------ emission kernel

struct OutCol {
    uint8_t* valid;
    void* data;
    int32_t type;
    int32_t pad;
};

struct EmitAgg {
    int32_t kind, flags, op_code, off;
    int32_t nulls_off, in_type, out0, bp;
    int32_t bs, rp, rs, bit;
    int32_t off2;    // a merge's count buffer total (SUM_I64 words), -1: none
    int32_t ne_off;  // a merged decimal sum's non-empty buffer rows, -1: none
};

// EmitAgg.flags beyond vr_agg_flags: the handle has no grouping key.
enum { EF_UNGROUPED = 1 << 16 };

struct EmitArgs {
    TableView t;
    KeySpec ks;
    const uint64_t* order;
    int64_t from, m;
    EmitAgg aggs[VR_MAX_AGGS];
    int32_t naggs, flags_off;
    OutCol keys[VR_MAX_KEYS];
    OutCol outs[2 * VR_MAX_AGGS];
};

// Writes row o's bit of a bitmap: the wave's 64 bits are gathered with a ballot and every eighth
// lane stores one byte (rows of a wave are consecutive and start on a multiple of 64).
__device__ __forceinline__ void put_bit(uint8_t* bm, int64_t o, bool v, bool in_range) {
    uint64_t mask = __ballot(v && in_range);
    int lane = __lane_id();
    if (bm && in_range && (lane & 7) == 0) bm[o >> 3] = (uint8_t)(mask >> lane);
}

__device__ __forceinline__ void put_i64(const OutCol& c, int64_t o, bool ok, int64_t v, bool in) {
    put_bit(c.valid, o, ok, in);
    if (in) ((int64_t*)c.data)[o] = ok ? v : 0;
}
__device__ __forceinline__ void put_d128(const OutCol& c, int64_t o, bool ok, uint64_t lo, uint64_t hi, bool in) {
    put_bit(c.valid, o, ok, in);
    if (in) {
        ((uint64_t*)c.data)[2 * o] = ok ? lo : 0;
        ((uint64_t*)c.data)[2 * o + 1] = ok ? hi : 0;
    }
}

// t (192 bits, little-endian limbs) += v * 2^shift, v unsigned or sign-extended.
__device__ __forceinline__ void add_shifted(uint64_t* t, uint64_t v, bool is_signed, int shift) {
    uint64_t ext = (is_signed && (int64_t)v < 0) ? ~0ull : 0ull;
    uint64_t x[3] = {v, ext, ext};  // v as 192 bits
    if (shift >= 64) {
        x[2] = x[1];
        x[1] = x[0];
        x[0] = 0;
        shift -= 64;
    }
    if (shift) {
        x[2] = (x[2] << shift) | (x[1] >> (64 - shift));
        x[1] = (x[1] << shift) | (x[0] >> (64 - shift));
        x[0] <<= shift;
    }
    uint64_t c = 0;
    for (int j = 0; j < 3; j++) {
        uint64_t s1 = t[j] + x[j];
        uint64_t c1 = s1 < t[j] ? 1ull : 0ull;
        uint64_t s2 = s1 + c;
        c = c1 + (s2 < s1 ? 1ull : 0ull);
        t[j] = s2;
    }
}

// The exact total of a sum state as 192 bits; `fits` = it is a signed 128-bit value.
__device__ __forceinline__ void sum_total(const uint64_t* w, bool d128, uint64_t* lo, uint64_t* hi, bool* fits) {
    uint64_t t[3] = {0, 0, 0};
    if (d128) {
        add_shifted(t, w[0], false, 0);
        add_shifted(t, w[1], false, 32);
        add_shifted(t, w[2], false, 64);
        add_shifted(t, w[3], true, 96);
    } else {
        add_shifted(t, w[0], false, 0);
        add_shifted(t, w[1], true, 32);
    }
    *lo = t[0];
    *hi = t[1];
    *fits = t[2] == (uint64_t)((int64_t)t[1] >> 63);
}

// Emission: one output row per thread.  STAGED: the workgroup's state rows are first copied into
// LDS with whole rows read by neighbouring lanes (a wave reads 64 consecutive words of several
// rows at once instead of one word of 64 rows), padded to an odd pitch of 64-bit words.
constexpr int EMIT_BLOCK = 256;
constexpr int EMIT_STAGE_MAX_PITCH = 33;  // state rows of at most this many words are staged

__host__ __device__ inline int emit_pitch(int stride) { return stride | 1; }

template <bool STAGED>
__global__ __launch_bounds__(EMIT_BLOCK) void agg_emit_kernel(EmitArgs a) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = o < a.m;
    uint64_t slot = in ? a.order[a.from + o] : 0;
    const uint64_t* g;
    if (STAGED) {
        extern __shared__ uint64_t emit_rows[];
        __shared__ uint64_t emit_slots[EMIT_BLOCK];
        const int S = a.t.stride, pitch = emit_pitch(S);
        emit_slots[threadIdx.x] = slot;
        __syncthreads();
        const int64_t left = a.m - (int64_t)blockIdx.x * blockDim.x;
        const int nrows = left < EMIT_BLOCK ? (int)left : EMIT_BLOCK;
        for (int i = threadIdx.x; i < nrows * S; i += EMIT_BLOCK) {
            const int r = i / S, w = i - r * S;
            emit_rows[r * pitch + w] = a.t.state[emit_slots[r] * (uint64_t)S + w];
        }
        __syncthreads();
        g = emit_rows + threadIdx.x * pitch;
    } else {
        g = a.t.state + slot * (uint64_t)a.t.stride;
    }
    // ---- keys
    const KeySpec& ks = a.ks;
    const uint64_t* wide = a.t.cells + slot * (uint64_t)a.t.width;  // KM_WIDE: the key's words, one per cell word
    auto w = [&](int i) { return (uint32_t)wide[i]; };
    uint64_t key64 = 0;
    bool null64 = false;
    if (in) {
        if (ks.mode == KM_WIDE) {
        } else if (slot == a.t.slots) {
            key64 = EMPTY_KEY;
        } else if (slot == a.t.slots + 1) {
            null64 = true;
        } else {
            key64 = a.t.cells[slot];
        }
    }
    for (int c = 0; c < ks.nkeys; c++) {
        const OutCol& out = a.keys[c];
        bool isnull = false;
        uint64_t lo = 0, hi = 0;
        if (ks.mode == KM_I64) {
            isnull = null64;
            lo = key64;
        } else if (ks.mode == KM_I32) {
            isnull = key64 == (1ull << 32);
            lo = (uint32_t)key64;
        } else if (ks.mode == KM_2X32) {
            lo = c == 0 ? (key64 >> 32) : (uint32_t)key64;
        } else if (in) {
            isnull = ks.null_word >= 0 && ((w(ks.null_word) >> c) & 1u);
            int o2 = ks.word_off[c];
            switch (ks.types[c]) {
                case VR_INT64:
                case VR_FLOAT64: lo = (uint64_t)w(o2) | ((uint64_t)w(o2 + 1) << 32); break;
                case VR_DECIMAL128:
                    lo = (uint64_t)w(o2) | ((uint64_t)w(o2 + 1) << 32);
                    hi = (uint64_t)w(o2 + 2) | ((uint64_t)w(o2 + 3) << 32);
                    break;
                default: lo = w(o2);
            }
        }
        if (!in) isnull = true;
        put_bit(out.valid, o, !isnull, in);
        switch (ks.types[c]) {
            case VR_BOOL: put_bit((uint8_t*)out.data, o, !isnull && lo != 0, in); break;
            case VR_INT32:
            case VR_UTF8:
                if (in) ((int32_t*)out.data)[o] = isnull ? 0 : (int32_t)(uint32_t)lo;
                break;
            case VR_INT64:
            case VR_FLOAT64:
                if (in) ((uint64_t*)out.data)[o] = isnull ? 0 : lo;
                break;
            case VR_DECIMAL128:
                if (in) {
                    ((uint64_t*)out.data)[2 * o] = isnull ? 0 : lo;
                    ((uint64_t*)out.data)[2 * o + 1] = isnull ? 0 : hi;
                }
                break;
        }
    }
    // ---- aggregates
    uint64_t err = 0;
    const uint64_t rows = in ? g[0] : 0;
    for (int k = 0; k < a.naggs; k++) {
        const EmitAgg& e = a.aggs[k];
        const OutCol& o0 = a.outs[e.out0];
        const OutCol& o1 = a.outs[e.out0 + 1];
        const bool result = e.flags & VR_AGG_RESULT;
        const bool ansi = e.flags & VR_AGG_CHECKED;
        const uint64_t* s = g + e.off;
        uint64_t cnt = rows - ((in && e.nulls_off >= 0) ? g[e.nulls_off] : 0);  // non-null rows
        switch (e.kind) {
            case VR_AGG_COUNT_STAR: put_i64(o0, o, in, (int64_t)rows, in); break;
            case VR_AGG_COUNT: put_i64(o0, o, in, (int64_t)cnt, in); break;
            case VR_AGG_SUM_LONG: {
                uint64_t lo = 0, hi = 0;
                bool fits = true;
                if (in) sum_total(s, false, &lo, &hi, &fits);
                bool long_fits = fits && hi == (uint64_t)((int64_t)lo >> 63);
                if (ansi && cnt && !long_fits) err |= ERR_LONG_OVERFLOW;
                put_i64(o0, o, cnt > 0, (int64_t)lo, in);
                break;
            }
            case VR_AGG_SUM_DOUBLE: put_i64(o0, o, cnt > 0, in ? (int64_t)s[0] : 0, in); break;
            case VR_AGG_AVG_DOUBLE: {
                double sum = in ? __longlong_as_double((long long)s[0]) : 0.0;
                if (result) {
                    double q = cnt ? sum / (double)cnt : 0.0;
                    put_i64(o0, o, cnt > 0, __double_as_longlong(q), in);
                } else {
                    put_i64(o0, o, true, __double_as_longlong(sum), in);
                    put_i64(o1, o, true, (int64_t)cnt, in);
                }
                break;
            }
            case VR_AGG_MERGE_COUNT: {
                uint64_t lo = 0, hi = 0;
                bool fits = true;
                if (in) sum_total(s, false, &lo, &hi, &fits);
                put_i64(o0, o, in, (int64_t)lo, in);  // LongSum: wraps like Java's long adds
                break;
            }
            case VR_AGG_MERGE_AVG_DOUBLE: {
                double sum = in ? __longlong_as_double((long long)s[0]) : 0.0;
                uint64_t c = 0, chi = 0;
                bool cf = true;
                if (in) sum_total(g + e.off2, false, &c, &chi, &cf);
                if (result) {
                    double q = c ? sum / (double)(int64_t)c : 0.0;
                    put_i64(o0, o, c != 0, __double_as_longlong(q), in);
                } else {
                    put_i64(o0, o, true, __double_as_longlong(sum), in);
                    put_i64(o1, o, true, (int64_t)c, in);
                }
                break;
            }
            case VR_AGG_MERGE_SUM_WIDE: {
                uint64_t lo = 0, hi = 0;
                bool fits = true;
                if (in) sum_total(s, e.op_code == OP_SUM_D128, &lo, &hi, &fits);
                const bool nonempty = in && g[e.ne_off] != 0;
                const bool poisoned = !fits || (in && g[e.nulls_off] != 0);
                const bool past = past_precision(lo, hi, e.bp);
                if (result) {
                    // If(isEmpty, null, CheckOverflowInSum(sum)) over the merged buffer
                    if (nonempty && ansi) err |= poisoned ? ERR_SUM_OF_OVERFLOW : (past ? ERR_DECIMAL_OVERFLOW : 0);
                    put_d128(o0, o, nonempty && !poisoned && !past, lo, hi, in);
                } else {
                    // zero while empty, null once poisoned or past the buffer precision
                    put_d128(o0, o, !nonempty || (!poisoned && !past), nonempty ? lo : 0, nonempty ? hi : 0, in);
                    put_bit(o1.valid, o, true, in);
                    put_bit((uint8_t*)o1.data, o, !nonempty, in);
                }
                break;
            }
            case VR_AGG_SUM_WIDE:
            case VR_AGG_AVG_WIDE:
            case VR_AGG_MERGE_AVG_WIDE: {
                uint64_t lo = 0, hi = 0;
                bool fits = true;
                if (in) sum_total(s, e.op_code == OP_SUM_D128, &lo, &hi, &fits);
                bool over = !fits || past_precision(lo, hi, e.bp);
                if (e.kind == VR_AGG_MERGE_AVG_WIDE) {
                    // The count is the partial counts' total, and a null partial sum poisons the group.
                    uint64_t chi = 0;
                    bool cf = true;
                    cnt = 0;
                    if (in) sum_total(g + e.off2, false, &cnt, &chi, &cf);
                    const bool poisoned = in && g[e.nulls_off] != 0;
                    if (!result) {
                        put_d128(o0, o, !poisoned && !over, lo, hi, in);
                        put_i64(o1, o, true, (int64_t)cnt, in);
                        break;
                    }
                    if (poisoned) {
                        if (cnt && ansi) err |= ERR_AVG_OF_OVERFLOW;
                        if (e.rp > 18) put_d128(o0, o, false, 0, 0, in);
                        else put_i64(o0, o, false, 0, in);
                        break;
                    }
                }
                if (result && (e.flags & EF_UNGROUPED) && e.kind != VR_AGG_SUM_WIDE) {
                    // An ungrouped final average divides the total as it is: Spark's generated code
                    // keeps it in a local that nothing checks against the buffer precision
                    // (WideDecimalAvgAgg.sumValue, VectorAggregates.scala:455-467, and the merge's
                    // exact ungrouped final, :583-604).  A total past 128 bits is past any result
                    // precision in update mode; the merge's exact total beyond 128 bits is left to
                    // the CPU kernels.
                    bool ok = cnt > 0;
                    uint64_t qlo = 0, qhi = 0;
                    if (ok && !fits) {
                        if (e.kind == VR_AGG_MERGE_AVG_WIDE) err |= ERR_WIDE_TOTAL;
                        else if (ansi) err |= ERR_DECIMAL_OVERFLOW;
                        ok = false;
                    }
                    if (ok) {
                        bool q_fits = decimal_avg(lo, hi, cnt, e.rs - e.bs, &qlo, &qhi);
                        if (!q_fits || past_precision(qlo, qhi, e.rp)) {
                            if (ansi) err |= ERR_DECIMAL_OVERFLOW;
                            ok = false;
                        }
                    }
                    if (e.rp > 18) put_d128(o0, o, ok, qlo, qhi, in);
                    else put_i64(o0, o, ok, (int64_t)qlo, in);
                    break;
                }
                if (e.kind == VR_AGG_SUM_WIDE) {
                    if (result) {
                        if (cnt && over && ansi) err |= ERR_DECIMAL_OVERFLOW;
                        put_d128(o0, o, cnt && !over, lo, hi, in);
                    } else {
                        put_d128(o0, o, cnt == 0 || !over, cnt ? lo : 0, cnt ? hi : 0, in);
                        put_bit(o1.valid, o, true, in);
                        put_bit((uint8_t*)o1.data, o, cnt == 0, in);
                    }
                } else if (!result) {
                    put_d128(o0, o, cnt == 0 || !over, cnt ? lo : 0, cnt ? hi : 0, in);
                    put_i64(o1, o, true, (int64_t)cnt, in);
                } else {
                    bool ok = cnt > 0;
                    uint64_t qlo = 0, qhi = 0;
                    if (ok && over) {
                        if (ansi) err |= ERR_AVG_OF_OVERFLOW;
                        ok = false;
                    }
                    if (ok) {
                        bool q_fits = decimal_avg(lo, hi, cnt, e.rs - e.bs, &qlo, &qhi);
                        if (!q_fits || past_precision(qlo, qhi, e.rp)) {
                            if (ansi) err |= ERR_DECIMAL_OVERFLOW;
                            ok = false;
                        }
                    }
                    if (e.rp > 18) put_d128(o0, o, ok, qlo, qhi, in);
                    else put_i64(o0, o, ok, (int64_t)qlo, in);
                }
                break;
            }
            case VR_AGG_MIN:
            case VR_AGG_MAX: {
                bool isMin = e.kind == VR_AGG_MIN;
                if (e.in_type == VR_INT32) {
                    int64_t v = in ? (int64_t)s[0] : 0;
                    bool has = in && v != (isMin ? INT64_MAX : INT64_MIN);
                    put_bit(o0.valid, o, has, in);
                    if (in) ((int32_t*)o0.data)[o] = has ? (int32_t)v : 0;
                } else if (e.in_type == VR_INT64) {
                    bool has = in && ((g[a.flags_off] >> e.bit) & 1ull);
                    put_i64(o0, o, has, has ? (int64_t)(s[0] ^ 0x8000000000000000ull) : 0, in);
                } else if (e.in_type == VR_DECIMAL128) {
                    const bool has = in && s[2] != 0;  // Decimal128MinMax.writeTo: null without a value
                    put_d128(o0, o, has, has ? s[0] : 0, has ? s[1] : 0, in);
                } else {
                    uint64_t code = in ? s[0] : 0;
                    bool has = in && code != (isMin ? ~0ull : 0ull);
                    uint64_t bits = unorder_f64(code);
                    if (has && bits == 0) bits = (s[1] & 1ull) << 63;  // the first-seen zero's sign
                    else if (has && is_nan_bits(bits)) bits = s[3];    // the first-seen NaN's payload
                    put_i64(o0, o, has, (int64_t)bits, in);
                }
                break;
            }
        }
    }
    if (err) atomicOr((unsigned long long*)&a.t.counters[C_ERROR], (unsigned long long)err);
}

// ------------------------------------------------------------------ host side

static uint64_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

static int key_words(int32_t type) {
    switch (type) {
        case VR_INT64:
        case VR_FLOAT64: return 2;
        case VR_DECIMAL128: return 4;
        default: return 1;
    }
}

// A UTF8 key column's dictionary on the device, and the code -> id map of the last input
// dictionary mapped through it (kept while the same dictionary keeps arriving).
struct StrKey {
    StrTable table;
    DevBuf remap;
    int32_t remap_len = 0;
    const void* last_dict = nullptr;
    int64_t last_len = -1;
    const int32_t* last_offsets = nullptr;
};

struct AggImpl {
    Device* dev = nullptr;
    std::vector<int32_t> key_types, value_types;
    std::vector<uint8_t> key_nullable;
    std::vector<vr_agg_desc> aggs;
    bool det = false;

    UpdateArgs proto{};  // ops, key spec, identity row
    int stride = 0;
    std::vector<EmitAgg> emits;
    std::vector<int32_t> out_types;

    // table: 2^pbits regions of slots >> pbits slots each (rgroups: groups per region)
    DevBuf cells, state, counters, rgroups;
    uint64_t slots = 0;
    int pbits = 0;
    double max_load = 0.5;
    int64_t groups_upper = 0;  // an upper bound of the groups (exact after a sync)
    int64_t rows_seen = 0;
    int32_t epoch = 0;
    bool has_f64_minmax = false;
    // Every update records its rows' slots and folds some states in row order after the kernel
    // (ordered_fold_kernel): the deterministic mode, or a 128-bit MIN/MAX.
    bool ordered = false;
    // Integer and decimal sums are kept as sums of 32-bit pieces in 64-bit words: exact until one
    // group has folded 2^32 rows.  The handle refuses an update past that many rows since its
    // creation or last reset (the caller emits and resets, or runs the CPU kernels).
    bool half_sums = false;
    uint64_t rows_since_reset = 0;

    std::vector<StrKey> strs;  // UTF8 keys: each one's device dictionary
    Staging staging;
    DevBuf ids_stage;  // host-mapped UTF8 ids (pinned copies go through `staging`)

    // emission
    DevBuf order, order_tmp, firsts, firsts_sorted, cub_tmp, out_stage;
    int64_t order_count = -1;

    // deterministic fold
    DevBuf row_slot, sort_keys, sort_vals, sort_vals_in;

    // batched launches
    DevBuf batch_descs;
    HostBuf desc_host, tile_host;  // the partitioned update's batch descriptors and tile -> batch map
    DevBuf resume;  // guarded launches: the tile each workgroup resumes at

    // partitioned update: hot-set sampling, record counts and ranges, record pool, merge resume points
    DevBuf p_cand, p_cand_count, p_hist, p_hot_keys, p_hot_part, p_hot_total;
    DevBuf p_pool, p_count, p_base, p_rtot, p_rstart, p_resume, p_tile_batch, p_pending;
    // Ungrouped aggregation (no grouping key: AggKernels' reductions): one group, keyed by a
    // constant INT64 column of zeros the library supplies, present from creation on so that an
    // empty input still emits its one row.
    bool ungrouped = false;
    DevBuf zero_keys;
    int64_t zero_rows = 0;
    int part_min_rows = 1 << 22;  // inputs from this many rows take the partitioned update

    // blocks_per_cu and lds_entries 0: the defaults of the call's path (launch_shape)
    vr_launch_config cfg{256, 4, 0, -1, 0, 4, 0, 0};
    int lds_budget = 64 * 1024;  // LDS bytes per workgroup for the hot-key cache

    // the stream of the last work queued on the handle, and its completion event: work queued on
    // another stream (or a host read) first waits for it
    hipStream_t last_stream = nullptr;
    hipEvent_t last_ev = nullptr;
    // measurement aid: events around every update-kernel launch while `timing` is on
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timer_events;
    size_t timer_used = 0;
    AggImpl() = default;
    AggImpl(const AggImpl&) = delete;
    AggImpl& operator=(const AggImpl&) = delete;
    ~AggImpl() {
        if (last_ev) (void)hipEventDestroy(last_ev);
        for (auto& e : timer_events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
};

static vr_status timer_begin(AggImpl& h, hipStream_t s) {
    if (!h.timing) return VR_OK;
    if (h.timer_used == h.timer_events.size()) {
        std::pair<hipEvent_t, hipEvent_t> e{};
        VR_TRY_HIP(hipEventCreate(&e.first));
        VR_TRY_HIP(hipEventCreate(&e.second));
        h.timer_events.push_back(e);
    }
    VR_TRY_HIP(hipEventRecord(h.timer_events[h.timer_used].first, s));
    return VR_OK;
}

static vr_status timer_end(AggImpl& h, hipStream_t s) {
    if (!h.timing) return VR_OK;
    VR_TRY_HIP(hipEventRecord(h.timer_events[h.timer_used++].second, s));
    return VR_OK;
}

static vr_status compile_spec(AggImpl& h) {
    UpdateArgs& p = h.proto;
    std::memset(&p, 0, sizeof(p));
    const int nk = (int)h.key_types.size();
    if (nk < 1 || nk > VR_MAX_KEYS) return fail(VR_ERR_UNSUPPORTED, "between 1 and 8 grouping keys are supported");
    if ((int)h.value_types.size() > VR_MAX_VALUES) return fail(VR_ERR_UNSUPPORTED, "at most 16 value columns");
    KeySpec& ks = p.ks;
    ks.nkeys = nk;
    bool any_nullable = false;
    int words = 0, narrow = 0;
    for (int c = 0; c < nk; c++) {
        int32_t t = h.key_types[c];
        if (t < VR_BOOL || t > VR_DECIMAL128) return fail(VR_ERR_UNSUPPORTED, "key lane type " + std::to_string(t));
        ks.types[c] = t;
        ks.word_off[c] = words;
        words += key_words(t);
        if (key_words(t) == 1) narrow++;
        if (h.key_nullable[c]) any_nullable = true;
    }
    if (nk == 1 && (h.key_types[0] == VR_INT64 || h.key_types[0] == VR_FLOAT64)) ks.mode = KM_I64;
    else if (nk == 1 && narrow == 1) ks.mode = KM_I32;
    else if (nk == 2 && narrow == 2 && !any_nullable) ks.mode = KM_2X32;
    else ks.mode = KM_WIDE;
    ks.null_word = -1;
    if (ks.mode == KM_WIDE) {
        if (any_nullable) ks.null_word = words++;
        ks.width = words;
    } else {
        ks.width = 1;
    }
    // operations, shared where two aggregates need the same state; ROWS is word 0 of every row
    std::vector<Op> ops;
    int off = 0;
    std::vector<uint64_t> ident;
    std::vector<uint8_t> comb;  // how two partial states combine, word by word
    auto add_op = [&](int32_t code, int32_t col, int nwords, int32_t gate = -1) -> int {
        for (size_t k = 0; k < ops.size(); k++)
            if (ops[k].code == code && ops[k].col == col && ops[k].gate == gate) return (int)k;
        Op op{};
        op.code = code;
        op.col = col;
        op.off = off;
        op.gate = (int16_t)gate;
        ops.push_back(op);
        for (int i = 0; i < nwords; i++) {
            ident.push_back(0);
            comb.push_back(CB_ADD);
        }
        switch (code) {
            case OP_SUM_F64: comb[off] = CB_FADD; break;
            case OP_MIN_I32:
                ident[off] = (uint64_t)INT64_MAX;
                comb[off] = CB_SMIN;
                break;
            case OP_MAX_I32:
                ident[off] = (uint64_t)INT64_MIN;
                comb[off] = CB_SMAX;
                break;
            case OP_MIN_I64:
                ident[off] = ~0ull;
                comb[off] = CB_UMIN;
                break;
            case OP_MAX_I64: comb[off] = CB_UMAX; break;
            case OP_MIN_F64:
            case OP_MAX_F64:
                // order code; first zero row; first NaN row; the NaN's bits (set on the final table)
                if (code == OP_MIN_F64) ident[off] = ~0ull;
                ident[off + 1] = ~0ull;
                ident[off + 2] = ~0ull;
                comb[off] = code == OP_MIN_F64 ? CB_UMIN : CB_UMAX;
                comb[off + 1] = CB_UMIN;
                comb[off + 2] = CB_UMIN;
                comb[off + 3] = CB_OR;
                break;
            case OP_FIRST_ROW:
                ident[off] = ~0ull;
                comb[off] = CB_UMIN;
                break;
        }
        off += nwords;
        return (int)ops.size() - 1;
    };
    add_op(OP_ROWS, -1, 1);
    int min_max_i64 = 0;
    uint32_t checked_sums = 0;
    h.emits.clear();
    h.out_types.clear();
    for (auto& d : h.aggs) {
        EmitAgg e{};
        e.kind = d.kind;
        e.flags = d.flags;
        e.bp = d.buffer_precision;
        e.bs = d.buffer_scale;
        e.rp = d.result_precision;
        e.rs = d.result_scale;
        e.out0 = (int)h.out_types.size();
        e.nulls_off = e.off2 = e.ne_off = -1;
        if (h.ungrouped) e.flags |= EF_UNGROUPED;
        int32_t in_type = -1, in2_type = -1;
        if (d.kind != VR_AGG_COUNT_STAR) {
            if (d.input < 0 || d.input >= (int)h.value_types.size())
                return fail(VR_ERR_INVALID_ARGUMENT, "aggregate input column out of range");
            in_type = h.value_types[d.input];
        }
        const bool two = d.kind == VR_AGG_MERGE_SUM_WIDE || d.kind == VR_AGG_MERGE_AVG_WIDE || d.kind == VR_AGG_MERGE_AVG_DOUBLE;
        if (two) {
            if (d.input2 < 0 || d.input2 >= (int)h.value_types.size())
                return fail(VR_ERR_INVALID_ARGUMENT, "second buffer column of a merge out of range");
            in2_type = h.value_types[d.input2];
            if (in2_type != (d.kind == VR_AGG_MERGE_SUM_WIDE ? VR_BOOL : VR_INT64))
                return fail(VR_ERR_INVALID_ARGUMENT, d.kind == VR_AGG_MERGE_SUM_WIDE ? "isEmpty buffer is not a boolean lane"
                                                                                      : "count buffer is not a bigint lane");
        }
        // Merge modes whose state and result are the update mode's over the buffer column
        // (VectorAggregates.compileMerge: SumLongAgg, SumDoubleAgg and MinMaxAgg over the buffer,
        // VectorAggregates.scala:1183-1200,1209-1210).
        if (d.kind == VR_AGG_MERGE_SUM_LONG) {
            if (in_type != VR_INT64) return fail(VR_ERR_INVALID_ARGUMENT, "sum buffer is not a bigint lane");
            e.kind = VR_AGG_SUM_LONG;
        } else if (d.kind == VR_AGG_MERGE_SUM_DOUBLE) {
            e.kind = VR_AGG_SUM_DOUBLE;
        } else if (d.kind == VR_AGG_MERGE_MIN || d.kind == VR_AGG_MERGE_MAX) {
            e.kind = d.kind == VR_AGG_MERGE_MIN ? VR_AGG_MIN : VR_AGG_MAX;
        }
        e.in_type = in_type;
        bool result = d.flags & VR_AGG_RESULT;
        int k = 0;
        auto nulls = [&]() { e.nulls_off = ops[add_op(OP_NULLS, d.input, 1)].off; };
        auto wide_sum = [&](int32_t gate) -> bool {
            if (in_type == VR_INT64) k = add_op(OP_SUM_I64, d.input, 2, gate);
            else if (in_type == VR_DECIMAL128) k = add_op(OP_SUM_D128, d.input, 4, gate);
            else return false;
            return true;
        };
        switch (e.kind) {
            case VR_AGG_COUNT_STAR: h.out_types.push_back(VR_INT64); break;
            case VR_AGG_COUNT:
                nulls();
                h.out_types.push_back(VR_INT64);
                break;
            case VR_AGG_SUM_LONG:
                if (in_type != VR_INT32 && in_type != VR_INT64) return fail(VR_ERR_UNSUPPORTED, "sum over a non-integral lane");
                k = add_op(OP_SUM_I64, d.input, 2);
                if (h.det && (d.flags & VR_AGG_CHECKED) && k < 32) checked_sums |= 1u << k;
                nulls();
                h.out_types.push_back(VR_INT64);
                break;
            case VR_AGG_SUM_DOUBLE:
            case VR_AGG_AVG_DOUBLE:
                if (in_type != VR_FLOAT64) return fail(VR_ERR_UNSUPPORTED, "double sum over a non-double lane");
                k = add_op(OP_SUM_F64, d.input, 1);
                nulls();
                h.out_types.push_back(VR_FLOAT64);
                if (e.kind == VR_AGG_AVG_DOUBLE && !result) h.out_types.push_back(VR_INT64);
                break;
            case VR_AGG_SUM_WIDE:
            case VR_AGG_AVG_WIDE:
                if (in_type == VR_INT32 || in_type == VR_INT64) k = add_op(OP_SUM_I64, d.input, 2);
                else if (in_type == VR_DECIMAL128) k = add_op(OP_SUM_D128, d.input, 4);
                else return fail(VR_ERR_UNSUPPORTED, "decimal sum over a non-integral lane");
                nulls();
                if (e.kind == VR_AGG_SUM_WIDE) {
                    h.out_types.push_back(VR_DECIMAL128);
                    if (!result) h.out_types.push_back(VR_BOOL);
                } else if (result) {
                    if (d.result_scale < d.buffer_scale) return fail(VR_ERR_UNSUPPORTED, "result scale below the buffer scale");
                    h.out_types.push_back(d.result_precision > 18 ? VR_DECIMAL128 : VR_INT64);
                } else {
                    h.out_types.push_back(VR_DECIMAL128);
                    h.out_types.push_back(VR_INT64);
                }
                break;
            case VR_AGG_MIN:
            case VR_AGG_MAX: {
                bool mn = e.kind == VR_AGG_MIN;
                if (in_type == VR_INT32) k = add_op(mn ? OP_MIN_I32 : OP_MAX_I32, d.input, 1);
                else if (in_type == VR_INT64) {
                    size_t before = ops.size();
                    k = add_op(mn ? OP_MIN_I64 : OP_MAX_I64, d.input, 1);
                    if (ops.size() != before) {
                        if (min_max_i64 >= 64) return fail(VR_ERR_UNSUPPORTED, "at most 64 bigint min/max");
                        ops[k].bit = min_max_i64++;
                    }
                } else if (in_type == VR_FLOAT64) {
                    k = add_op(mn ? OP_MIN_F64 : OP_MAX_F64, d.input, 4);
                    h.has_f64_minmax = true;
                } else if (in_type == VR_DECIMAL128) {
                    k = add_op(mn ? OP_MIN_D128 : OP_MAX_D128, d.input, 3);
                    h.ordered = true;
                } else {
                    return fail(VR_ERR_UNSUPPORTED, "min/max over this lane runs on the CPU kernels");
                }
                h.out_types.push_back(in_type);
                break;
            }
            // COUNT's merge: a LongSum over the partial counts, never null (CountMergeAgg,
            // VectorAggregates.scala:1042-1062).
            case VR_AGG_MERGE_COUNT:
                if (in_type != VR_INT64) return fail(VR_ERR_INVALID_ARGUMENT, "count buffer is not a bigint lane");
                k = add_op(OP_SUM_I64, d.input, 2);
                h.out_types.push_back(VR_INT64);
                break;
            // AVG's merge: the sums (DoubleSum, strict in deterministic mode) and the counts added
            // component-wise (AverageMergeAgg, VectorAggregates.scala:1065-1090).
            case VR_AGG_MERGE_AVG_DOUBLE:
                if (in_type != VR_FLOAT64) return fail(VR_ERR_INVALID_ARGUMENT, "avg sum buffer is not a double lane");
                k = add_op(OP_SUM_F64, d.input, 1);
                e.off2 = ops[add_op(OP_SUM_I64, d.input2, 2)].off;
                h.out_types.push_back(VR_FLOAT64);
                if (!result) h.out_types.push_back(VR_INT64);
                break;
            // The wide decimal sum's merge: buffer rows whose isEmpty is false (and non-null) mark the
            // group non-empty; a null sum on such a row poisons it; the others add exactly
            // (WideDecimalSumMergeAgg.State.merge, VectorAggregates.scala:688-712).
            case VR_AGG_MERGE_SUM_WIDE:
                if (!wide_sum(d.input2)) return fail(VR_ERR_INVALID_ARGUMENT, "decimal sum buffer is not a decimal lane");
                e.ne_off = ops[add_op(OP_ROWS, -1, 1, d.input2)].off;
                e.nulls_off = ops[add_op(OP_NULLS, d.input, 1, d.input2)].off;
                h.out_types.push_back(VR_DECIMAL128);
                if (!result) h.out_types.push_back(VR_BOOL);
                break;
            // The wide decimal average's merge: counts added, a null sum poisons the group, the sums
            // add exactly (WideDecimalAvgMergeAgg.State.merge, VectorAggregates.scala:562-578).
            case VR_AGG_MERGE_AVG_WIDE:
                if (!wide_sum(-1)) return fail(VR_ERR_INVALID_ARGUMENT, "decimal avg buffer is not a decimal lane");
                nulls();
                e.off2 = ops[add_op(OP_SUM_I64, d.input2, 2)].off;
                if (result) {
                    if (d.result_scale < d.buffer_scale) return fail(VR_ERR_UNSUPPORTED, "result scale below the buffer scale");
                    h.out_types.push_back(d.result_precision > 18 ? VR_DECIMAL128 : VR_INT64);
                } else {
                    h.out_types.push_back(VR_DECIMAL128);
                    h.out_types.push_back(VR_INT64);
                }
                break;
            default:
                return fail(VR_ERR_UNSUPPORTED, "aggregate kind " + std::to_string(d.kind) + " runs on the CPU kernels");
        }
        e.op_code = ops[k].code;
        e.off = ops[k].off;
        e.bit = ops[k].bit;
        h.emits.push_back(e);
    }
    if ((int)h.emits.size() > VR_MAX_AGGS) return fail(VR_ERR_UNSUPPORTED, "at most 24 aggregates");
    if (h.det) add_op(OP_FIRST_ROW, -1, 1);
    p.flags_off = -1;
    if (min_max_i64 > 0) {
        p.flags_off = off;
        ident.push_back(0);
        comb.push_back(CB_OR);
        off++;
    }
    if ((int)ops.size() > VR_MAX_OPS || off > 64) return fail(VR_ERR_UNSUPPORTED, "aggregate state too wide");
    for (size_t k = 0; k < ops.size(); k++) p.ops[k] = ops[k];
    p.nops = (int)ops.size();
    for (int i = 0; i < off; i++) {
        p.ident[i] = ident[i];
        p.comb[i] = comb[i];
    }
    h.stride = off;
    p.det = h.det ? 1 : 0;
    p.checked_sums = checked_sums;
    p.filter_col = -1;
    return VR_OK;
}

static int log2_of(uint64_t v) {
    int l = 0;
    while ((1ull << l) < v) l++;
    return l;
}

// Regions of at least 2048 slots, at most 1024 of them.
static int choose_pbits(uint64_t slots) { return std::max(0, std::min(10, log2_of(slots) - 11)); }

static TableView view_of(AggImpl& h) {
    TableView t{};
    t.cells = h.cells.as<uint64_t>();
    t.state = h.state.as<uint64_t>();
    t.counters = h.counters.as<uint64_t>();
    // groups per region are read by the partitioned merge only, which wide keys never take: their
    // inserts skip the count (one device atomic per new group on a few hundred counters)
    t.rgroups = h.proto.ks.mode == KM_WIDE ? nullptr : h.rgroups.as<uint32_t>();
    t.slots = h.slots;
    t.mask = h.slots - 1;
    t.pbits = h.pbits;
    t.rbits = log2_of(h.slots) - h.pbits;
    t.rmask = (1ull << t.rbits) - 1;
    t.width = h.proto.ks.width;
    t.stride = h.stride;
    return t;
}

static int grid_for(uint64_t work, int block, int cap) {
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (uint64_t)cap) g = (uint64_t)cap;
    return (int)g;
}

// Zeroes the groups-per-region counts of a table with 2^pbits regions.
static vr_status clear_rgroups(AggImpl& h, DevBuf& rgroups, int pbits, hipStream_t s) {
    VR_TRY(rgroups.ensure(h.dev, sizeof(uint32_t) << pbits));
    const uint64_t words = ((1ull << pbits) + 1) / 2;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(words, 256, 64)), dim3(256), 0, s, rgroups.as<uint64_t>(), words, 0ull);
    VR_TRY_HIP(hipGetLastError());
    return VR_OK;
}

// Allocates an empty table of `slots` slots in 2^pbits regions (cells empty, state at identity,
// region counts zero) into the given buffers.  State rows: the slots, two reserved rows, and the
// row that takes the rows of a full region.
static vr_status alloc_table(AggImpl& h, uint64_t slots, int pbits, DevBuf& cells, DevBuf& state, DevBuf& rgroups,
                             hipStream_t s) {
    const int W = h.proto.ks.width;
    VR_TRY(cells.ensure(h.dev, slots * W * sizeof(uint64_t)));
    VR_TRY(state.ensure(h.dev, (slots + 3) * h.stride * sizeof(uint64_t)));
    uint64_t empty = h.proto.ks.mode == KM_WIDE ? 0ull : EMPTY_KEY;
    int cap = h.dev->num_cus * 16;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(slots * W, 256, cap)), dim3(256), 0, s, cells.as<uint64_t>(),
                       (uint64_t)(slots * W), empty);
    UpdateArgs a = h.proto;
    a.t.stride = h.stride;
    hipLaunchKernelGGL(init_state_kernel, dim3(grid_for((slots + 3 + 63) / 64 * 64, 256, cap)), dim3(256), 0, s,
                       state.as<uint64_t>(), (uint64_t)(slots + 3), a);
    VR_TRY_HIP(hipGetLastError());
    VR_TRY(clear_rgroups(h, rgroups, pbits, s));
    return VR_OK;
}

static vr_status read_counters(AggImpl& h, uint64_t* out, hipStream_t s) {
    VR_TRY_HIP(hipMemcpyAsync(out, h.counters.p, C_WORDS * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (out[C_ERROR] & ERR_TABLE_FULL)
        return fail(VR_ERR_OOM, "a hash table region is full (keys that do not spread over the table)");
    if (out[C_ERROR] & ERR_INTERNAL) return fail(VR_ERR_DEVICE, "the partitioned update's record counts disagree");
    if (out[C_ERROR] & ERR_ORDER) return fail(VR_ERR_DEVICE, "the table's occupied slots differ from its group count");
    return VR_OK;
}

static int64_t groups_of(const uint64_t* c) {
    return (int64_t)c[C_GROUPS] + ((c[C_RESERVED] & RES_SENTINEL) ? 1 : 0) + ((c[C_RESERVED] & RES_NULL) ? 1 : 0);
}

// Moves the groups into a new table of `want` slots in 2^pbits regions (pbits < 0: the default
// for the size).
static vr_status grow_table(AggImpl& h, uint64_t want, hipStream_t s, int pbits = -1) {
    if (want > (1ull << 31)) return fail(VR_ERR_UNSUPPORTED, "more than 2^30 groups run on the CPU kernels");
    if (pbits < 0) pbits = std::max(choose_pbits(want), std::min(h.pbits, log2_of(want) - 11));  // keep the regions
    DevBuf ncells, nstate, nrg;
    VR_TRY(alloc_table(h, want, pbits, ncells, nstate, nrg, s));
    TableView from = view_of(h);
    UpdateArgs a = h.proto;
    std::swap(h.cells.p, ncells.p);
    std::swap(h.cells.bytes, ncells.bytes);
    std::swap(h.state.p, nstate.p);
    std::swap(h.state.bytes, nstate.bytes);
    std::swap(h.rgroups.p, nrg.p);
    std::swap(h.rgroups.bytes, nrg.bytes);
    h.slots = want;
    h.pbits = pbits;
    a.t = view_of(h);
    hipLaunchKernelGGL(migrate_kernel, dim3(grid_for(from.slots + 2, 256, h.dev->num_cus * 16)), dim3(256), 0, s, from, a);
    VR_TRY_HIP(hipGetLastError());
    VR_TRY_HIP(hipStreamSynchronize(s));
    h.dev->count(&vr_stats::table_grows, 1);
    return VR_OK;
}

// Makes room for `n` more rows: grows (and migrates) the table when the groups could pass the
// load limit.  The group count is read back only when the cheap upper bound says it might.
static vr_status ensure_capacity(AggImpl& h, int64_t n, hipStream_t s) {
    if ((double)(h.groups_upper + n) <= h.max_load * (double)h.slots) return VR_OK;
    uint64_t c[C_WORDS];
    VR_TRY(read_counters(h, c, s));
    int64_t g = groups_of(c);
    h.groups_upper = g;
    if ((double)(g + n) <= h.max_load * (double)h.slots) return VR_OK;
    return grow_table(h, next_pow2((uint64_t)((double)(g + n) / h.max_load) + 1), s);
}

static vr_status init_table(AggImpl& h, int64_t capacity_hint) {
    uint64_t want = next_pow2((uint64_t)((double)std::max<int64_t>(capacity_hint, 1024) / h.max_load) + 1);
    if (want < 2048) want = 2048;
    hipStream_t s = h.dev->stream;
    h.pbits = choose_pbits(want);
    VR_TRY(alloc_table(h, want, h.pbits, h.cells, h.state, h.rgroups, s));
    VR_TRY(h.counters.ensure(h.dev, C_WORDS * sizeof(uint64_t)));
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h.counters.as<uint64_t>(), (uint64_t)C_WORDS, 0ull);
    VR_TRY_HIP(hipGetLastError());
    VR_TRY_HIP(hipStreamSynchronize(s));
    h.slots = want;
    return VR_OK;
}

// The ungrouped handle's one group (after creation and after each reset).
static vr_status insert_single_group(AggImpl& h, hipStream_t s) {
    hipLaunchKernelGGL(insert_key_kernel, dim3(1), dim3(64), 0, s, view_of(h), 0ull);
    VR_TRY_HIP(hipGetLastError());
    h.groups_upper = 1;
    return VR_OK;
}

// The ungrouped handle's key column for n rows: zeros on the device.
static vr_status zero_key_column(AggImpl& h, int64_t n, hipStream_t s, vr_column* col) {
    if (h.zero_rows < n) {
        const int64_t want = std::max<int64_t>(n, 65536);
        VR_TRY(h.zero_keys.ensure(h.dev, (size_t)want * sizeof(int64_t)));
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for((uint64_t)want, 256, h.dev->num_cus * 8)), dim3(256), 0, s,
                           h.zero_keys.as<uint64_t>(), (uint64_t)want, 0ull);
        VR_TRY_HIP(hipGetLastError());
        h.zero_rows = want;
    }
    std::memset(col, 0, sizeof(*col));
    col->type = VR_INT64;
    col->mem = VR_MEM_DEVICE;
    col->length = n;
    col->data = h.zero_keys.p;
    return VR_OK;
}

// ---- columns -> device descriptors (staging host buffers)

struct StagePlan {
    std::vector<StagePiece> pieces;
    size_t total = 0;
    void add(const void* src, size_t bytes, const void** dst) {
        if (!src || !bytes) return;
        pieces.push_back({src, bytes, dst});
        total = align_up(total, 256) + bytes;
    }
};

// The device id arrays of a call's plain UTF8 key columns: freed only once the stream's queued
// work, which reads them, has finished.
struct KeyIds {
    std::vector<DevBuf> bufs;
    hipStream_t s = nullptr;
    explicit KeyIds(hipStream_t stream) : s(stream) {}
    ~KeyIds() {
        if (!bufs.empty()) (void)hipStreamSynchronize(s);
    }
};

// Key c's ids for a UTF8 column, from the handle's device dictionary of that key
// (GroupKeyTable.toIds, GroupKeyTable.java:1308-1430): a dictionary-encoded column maps its
// dictionary's entries once per dictionary (the code -> id map stays on the device while the same
// dictionary keeps arriving); a plain column, in host or device memory, maps its rows into a
// device array of ids kept in `ids_store` until the update's kernels have run.
static vr_status map_utf8_key(AggImpl& h, int c, const vr_column& col, int64_t n, DCol& d, StagePlan& plan,
                              std::vector<DevBuf>& ids_store, hipStream_t s) {
    StrKey& sk = h.strs[c];
    if (col.dictionary) {
        const vr_column& dc = *col.dictionary;
        if (dc.mem != VR_MEM_HOST || dc.type != VR_UTF8 || !dc.offsets)
            return fail(VR_ERR_UNSUPPORTED, "a key dictionary must be a plain UTF8 column in host memory");
        if (sk.last_dict != dc.data || sk.last_len != dc.length || sk.last_offsets != dc.offsets) {
            VR_TRY_HIP(hipStreamSynchronize(s));  // a kernel may still read the old map
            VR_TRY(sk.remap.ensure(h.dev, (size_t)std::max<int64_t>(dc.length, 1) * sizeof(int32_t)));
            // a null entry maps to -1: the row is null through its dictionary
            VR_TRY(sk.table.map_host(h.dev, dc.offsets, (const uint8_t*)dc.data, dc.validity, dc.length, true, -1,
                                     sk.remap.as<int32_t>(), s));
            sk.remap_len = (int32_t)std::max<int64_t>(dc.length, 1);
            sk.last_dict = dc.data;
            sk.last_len = dc.length;
            sk.last_offsets = dc.offsets;
        }
        d.remap = sk.remap.as<int32_t>();
        d.remap_len = sk.remap_len;
        d.data = col.data;
        if (col.mem == VR_MEM_HOST) plan.add(col.data, (size_t)n * 4, &d.data);
        return VR_OK;
    }
    if (!col.offsets) return fail(VR_ERR_INVALID_ARGUMENT, "a plain UTF8 key column needs its offsets");
    ids_store.emplace_back();
    DevBuf& ids = ids_store.back();
    VR_TRY(ids.ensure(h.dev, (size_t)std::max<int64_t>(n, 1) * 4));
    if (col.mem == VR_MEM_HOST) {
        VR_TRY(sk.table.map_host(h.dev, col.offsets, (const uint8_t*)col.data, col.validity, n, true, 0, ids.as<int32_t>(), s));
    } else {
        const StrRows rows{col.offsets, (const uint8_t*)col.data, col.validity, n};
        VR_TRY(sk.table.map(h.dev, rows, true, 0, ids.as<int32_t>(), s));
    }
    d.remap = nullptr;
    d.data = ids.p;
    return VR_OK;
}

static vr_status describe(AggImpl& h, const vr_column* keys, const vr_column* values, int64_t n,
                          const uint8_t* selection, BatchDesc& b, StagePlan& plan, std::vector<DevBuf>& id_store,
                          int32_t* sel_mem, hipStream_t s) {
    std::memset(&b, 0, sizeof(b));
    b.n = n;
    int32_t mem = -1;
    auto note_mem = [&](int32_t m) -> vr_status {
        if (m != VR_MEM_HOST && m != VR_MEM_DEVICE) return fail(VR_ERR_INVALID_ARGUMENT, "bad vr_mem");
        return VR_OK;
    };
    const int nk = (int)h.key_types.size(), nv = (int)h.value_types.size();
    for (int c = 0; c < nk; c++) {
        const vr_column& col = keys[c];
        VR_TRY(note_mem(col.mem));
        if (col.type != h.key_types[c]) return fail(VR_ERR_INVALID_ARGUMENT, "key " + std::to_string(c) + " lane type differs from the spec");
        if (col.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "key column shorter than the batch");
        if (col.validity && !h.key_nullable[c]) return fail(VR_ERR_INVALID_ARGUMENT, "nulls in a key declared not nullable");
        DCol& d = b.keys[c];
        d.type = col.type;
        d.valid = col.validity;
        if (col.mem == VR_MEM_HOST) plan.add(col.validity, bitmap_bytes(n), (const void**)&d.valid);
        if (col.type == VR_UTF8) {
            VR_TRY(map_utf8_key(h, c, col, n, d, plan, id_store, s));
        } else {
            d.data = col.data;
            if (col.mem == VR_MEM_HOST) plan.add(col.data, lane_bytes(col.type, n), &d.data);
        }
        mem = col.mem;
    }
    for (int c = 0; c < nv; c++) {
        const vr_column& col = values[c];
        VR_TRY(note_mem(col.mem));
        if (col.type != h.value_types[c]) return fail(VR_ERR_INVALID_ARGUMENT, "value " + std::to_string(c) + " lane type differs from the spec");
        if (col.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "value column shorter than the batch");
        DCol& d = b.vals[c];
        d.type = col.type;
        d.valid = col.validity;
        d.data = col.data;
        if (col.mem == VR_MEM_HOST) {
            plan.add(col.validity, bitmap_bytes(n), (const void**)&d.valid);
            plan.add(col.data, lane_bytes(col.type, n), &d.data);
        }
        mem = col.mem;
    }
    b.sel = selection;
    if (selection && mem == VR_MEM_HOST) plan.add(selection, bitmap_bytes(n), (const void**)&b.sel);
    *sel_mem = mem;
    return VR_OK;
}

static vr_status stage(AggImpl& h, StagePlan& plan, hipStream_t s, int* slot) {
    return h.staging.stage(h.dev, plan.pieces, s, slot);
}

// Default workgroups per CU and LDS cache entries.  A batch set runs a persistent grid of 4 per CU
// with 512 entries each.  One batch runs 2 per CU with up to 1,024 entries: each workgroup then folds
// two or more tiles into a cache that holds a thousand groups before it flushes (one 1M-row call over
// 1,000 int64 keys: 0.27 ms instead of 0.33 ms; 3M rows: 0.45 instead of 0.67 ms), for 2-3% more
// time per call where nearly every row is a new group.
constexpr int BATCHES_BLOCKS_PER_CU = 4, BATCHES_LDS_ENTRIES = 512;
constexpr int SINGLE_BLOCKS_PER_CU = 2, SINGLE_LDS_ENTRIES = 1024;

static int batches_cap(const AggImpl& h, const vr_launch_config* cfg) {
    const int bpc = (cfg && cfg->blocks_per_cu) ? cfg->blocks_per_cu
                    : h.cfg.blocks_per_cu       ? h.cfg.blocks_per_cu
                                                : BATCHES_BLOCKS_PER_CU;
    return h.dev->num_cus * bpc;
}

static void launch_shape(AggImpl& h, const vr_launch_config* cfg, int64_t n, UpdateArgs& a, int* grid,
                         int* block, size_t* lds_bytes, bool multi) {
    vr_launch_config c = h.cfg;
    if (cfg) {
        if (cfg->rows_per_thread) c.rows_per_thread = cfg->rows_per_thread;
        if (cfg->blocks_per_cu) c.blocks_per_cu = cfg->blocks_per_cu;
        if (cfg->lds_preaggregate >= 0) c.lds_preaggregate = cfg->lds_preaggregate;
        if (cfg->lds_entries) c.lds_entries = cfg->lds_entries;
        if (cfg->flush_tiles) c.flush_tiles = cfg->flush_tiles;
    }
    *block = AGG_BLOCK;  // the kernel's launch bounds
    if (multi || (cfg && cfg->rows_per_thread)) {
        a.rows_per_thread = std::max(1, c.rows_per_thread);
    } else {
        // one batch: enough tiles to give every CU a couple of workgroups
        int64_t want_tiles = (int64_t)h.dev->num_cus * 2;
        int64_t rpt = n / ((int64_t)AGG_BLOCK * want_tiles);
        a.rows_per_thread = (int)std::max<int64_t>(1, std::min<int64_t>(rpt, c.rows_per_thread));
    }
    int64_t T = (int64_t)(*block) * a.rows_per_thread;
    int64_t tiles = multi ? a.total_tiles : (n + T - 1) / T;
    if (!c.blocks_per_cu) c.blocks_per_cu = multi ? BATCHES_BLOCKS_PER_CU : SINGLE_BLOCKS_PER_CU;
    if (!c.lds_entries) c.lds_entries = multi ? BATCHES_LDS_ENTRIES : SINGLE_LDS_ENTRIES;
    int cap = h.dev->num_cus * c.blocks_per_cu;
    *grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    bool lds = c.lds_preaggregate != 0;
    int E = lds ? std::max(64, c.lds_entries) : 0;
    if (!multi && lds) {
        // one batch: the cache a workgroup clears and flushes need not hold more entries than a
        // tile has rows (65,536-row batches: one tile of 256 rows per workgroup, 256 entries, a
        // fifth less time per call than 512)
        E = std::min<int64_t>(E, std::max<int64_t>(64, next_pow2((uint64_t)T)));
    }
    while (E & (E - 1)) E &= E - 1;  // a power of two
    if (E > 2048) E = 2048;
    size_t bytes = (size_t)E * (1 + h.stride) * sizeof(uint64_t);
    while (E > 64 && bytes > (size_t)h.lds_budget) {
        E >>= 1;
        bytes = (size_t)E * (1 + h.stride) * sizeof(uint64_t);
    }
    a.lds_entries = E;
    a.flush_tiles = std::max(1, c.flush_tiles);
    *lds_bytes = (size_t)E * (1 + h.stride) * sizeof(uint64_t);
}

// Inputs up to GUARD_ROWS rows are launched once the table has room for every row to be a new
// group (no read-back while the upper bound allows it); longer inputs take the guarded launch, so a
// long input over few groups keeps a table sized by its groups, not by its rows.
constexpr int64_t GUARD_ROWS = 1ll << 22;
// Share of the slots a guarded launch may fill: the load limit plus what the workgroups in flight
// can still insert after they see it.
constexpr double GUARD_FULL = 0.95;

// Runs the update over tiles [a.tile_lo, a.total_tiles) whatever the table's headroom: workgroups
// stop at the load limit, the table grows, and the same grid resumes where each one stopped.
template <bool MULTI>
static vr_status launch_guarded(AggImpl& h, UpdateArgs& a, int grid, int block, size_t lds_bytes, hipStream_t s,
                                int* launches) {
    const int64_t T = (int64_t)block * a.rows_per_thread;
    // Groups the table may gain beyond the limit: between two looks at the shared count a workgroup
    // inserts at most a tile's rows and its LDS entries, and the count it sees lags by as much.
    const int64_t margin = 2 * (int64_t)grid * (T + a.lds_entries);
    VR_TRY(h.resume.ensure(h.dev, (size_t)grid * sizeof(int64_t)));
    a.resume = h.resume.as<int64_t>();
    a.resumed = 0;
    while (true) {
        VR_TRY(ensure_capacity(h, margin, s));
        if ((GUARD_FULL - h.max_load) * (double)h.slots < (double)margin)
            VR_TRY(grow_table(h, next_pow2((uint64_t)((double)margin / (GUARD_FULL - h.max_load)) + 1), s));
        a.t = view_of(h);
        a.soft_limit = (uint64_t)(h.max_load * (double)h.slots);
        VR_TRY(timer_begin(h, s));
        hipLaunchKernelGGL(agg_update_kernel<MULTI>, dim3(grid), dim3(block), lds_bytes, s, a);
        VR_TRY(timer_end(h, s));
        VR_TRY_HIP(hipGetLastError());
        (*launches)++;
        uint64_t c[C_WORDS];
        VR_TRY(read_counters(h, c, s));
        h.groups_upper = groups_of(c);
        if (!c[C_STOP]) return VR_OK;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h.counters.as<uint64_t>() + C_STOP, 1ull, 0ull);
        VR_TRY_HIP(hipGetLastError());
        a.resumed = 1;
    }
}

// ---- partitioned update (keys of at most 64 bits; the kernels' header describes the scheme)

static int lane_bytes_of(int32_t type) {
    switch (type) {
        case VR_INT32: return 4;
        case VR_INT64:
        case VR_FLOAT64: return 8;
        case VR_DECIMAL128: return 16;
    }
    return 0;
}

// The record layout of this spec: key (words 0-1), 16-byte lanes, 8-byte lanes, the row index,
// 4-byte lanes, the validity word.  False when the partitioned update does not apply to the spec.
static bool record_layout(const AggImpl& h, RecLayout& L) {
    std::memset(&L, 0, sizeof(L));
    const int nv = (int)h.value_types.size();
    L.nvals = nv;
    for (int c = 0; c < VR_MAX_VALUES; c++) {
        L.val_off[c] = -1;
        L.val_type[c] = c < nv ? h.value_types[c] : 0;
    }
    bool need_row = false;
    std::vector<int> carry(nv, 0);
    for (int k = 0; k < h.proto.nops; k++) {
        const Op& op = h.proto.ops[k];
        if (op.code == OP_FIRST_ROW) return false;
        if (op.gate >= 0) return false;  // records carry no gate column
        if (op.col < 0) continue;
        L.vmask |= 1u << op.col;
        if (op.code == OP_NULLS) continue;
        if (op.code == OP_MIN_F64 || op.code == OP_MAX_F64) need_row = true;
        carry[op.col] = lane_bytes_of(h.value_types[op.col]);
    }
    int w = 2;
    for (int size : {16, 8})
        for (int c = 0; c < nv; c++)
            if (carry[c] == size) {
                L.val_off[c] = w;
                w += size / 4;
            }
    L.row_off = -1;
    if (need_row) {
        L.row_off = w;
        w += 2;
    }
    for (int c = 0; c < nv; c++)
        if (carry[c] == 4) L.val_off[c] = w++;
    L.valid_off = w++;
    L.words = (w + 3) / 4 * 4;
    return L.words <= PART_REC_WORDS;
}

static_assert(sizeof(PartArgs) <= 4096, "kernel arguments are at most 4 KiB");

// The aggregates of each carried or counted value column as a program (ColProg).
static int column_programs(const AggImpl& h, const RecLayout& L, ColProg* prog) {
    int n = 0;
    for (int c = 0; c < L.nvals; c++) {
        if (!((L.vmask >> c) & 1u)) continue;
        ColProg cp;
        std::memset(&cp, 0, sizeof(cp));
        cp.col = (int16_t)c;
        cp.type = (int16_t)h.value_types[c];
        cp.rec_off = (int16_t)L.val_off[c];
        cp.nulls_off = cp.sum_off = cp.min_off = cp.max_off = -1;
        for (int k = 0; k < h.proto.nops; k++) {
            const Op& op = h.proto.ops[k];
            if (op.col != c) continue;
            switch (op.code) {
                case OP_NULLS: cp.flags |= CP_NULLS; cp.nulls_off = (int16_t)op.off; break;
                case OP_SUM_I64: cp.flags |= CP_SUM_I64; cp.sum_off = (int16_t)op.off; break;
                case OP_SUM_D128: cp.flags |= CP_SUM_D128; cp.sum_off = (int16_t)op.off; break;
                case OP_SUM_F64: cp.flags |= CP_SUM_F64; cp.sum_off = (int16_t)op.off; break;
                case OP_MIN_I32: cp.flags |= CP_MIN_I32; cp.min_off = (int16_t)op.off; break;
                case OP_MAX_I32: cp.flags |= CP_MAX_I32; cp.max_off = (int16_t)op.off; break;
                case OP_MIN_I64:
                    cp.flags |= CP_MIN_I64;
                    cp.min_off = (int16_t)op.off;
                    cp.min_bit = (int16_t)op.bit;
                    break;
                case OP_MAX_I64:
                    cp.flags |= CP_MAX_I64;
                    cp.max_off = (int16_t)op.off;
                    cp.max_bit = (int16_t)op.bit;
                    break;
                default: cp.flags |= CP_GENERIC; break;  // double MIN/MAX: first-seen zero and NaN rows
            }
        }
        // the lane is read when a record carries it or an aggregate other than a null count uses it
        if (cp.rec_off >= 0 || (cp.flags & ~CP_NULLS)) cp.width = (int16_t)lane_bytes_of(cp.type);
        prog[n++] = cp;
    }
    return n;
}

// LDS cache entries of the merge (0: the state rows are too wide for it).
static int merge_cache_entries(const AggImpl& h) {
    const int64_t budget = h.dev->lds_per_block - 2048;  // the kernel's static LDS (counters, claim bits)
    const int E = PB_FE * PB_BLOCK;
    if ((int64_t)merge_lds_bytes(E, h.stride) <= budget) return E;
    return 0;
}

// Record blocks of the pool for one partitioned update over `rows` rows: every active row may be
// a record, and each of the P regions' ranges ends in a partly filled block.
static uint64_t pool_blocks(int64_t rows, uint64_t P) { return (uint64_t)(rows + PART_BLK - 1) / PART_BLK + P; }

static bool partitioned_applies(const AggImpl& h, int64_t rows, RecLayout* L, const vr_launch_config* cfg) {
    const int32_t mode = (cfg && cfg->partitioned) ? cfg->partitioned : h.cfg.partitioned;
    if (mode < 0 || (mode == 0 && rows < h.part_min_rows)) return false;
    if (h.ordered || h.proto.ks.mode == KM_WIDE || rows <= 0) return false;
    // record positions are 32-bit
    if (pool_blocks(rows, 1ull << PART_MAX_PBITS) * PART_BLK >= (1ull << 32)) return false;
    if (!record_layout(h, *L)) return false;
    ColProg prog[VR_MAX_VALUES];
    if (column_programs(h, *L, prog) > PART_MAX_PROG) return false;
    return merge_cache_entries(h) > 0;
}

template <class K>
static vr_status allow_lds(K kernel, size_t bytes) {
    VR_TRY_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return VR_OK;
}

// Runs the partitioned update over the tiles [a.tile_lo, a.total_tiles) of the batches in `a`
// (tiles of PA_BLOCK * a.rows_per_thread rows, columns on the device).
// Launches a kernel compiled at run time with the partitioned update's arguments.
static vr_status launch_jit(hipFunction_t f, unsigned grid, unsigned block, size_t lds, hipStream_t s, PartArgs& pa) {
    void* args[] = {&pa};
    VR_TRY_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, (unsigned)lds, s, args, nullptr));
    return VR_OK;
}

static vr_status run_partitioned(AggImpl& h, UpdateArgs& a, int64_t rows, const RecLayout& L, hipStream_t s,
                                 int* launches, const PartJit* jit, bool* specialized) {
    const int S = h.stride;
    // Regions small enough for the merge cache to hold all the keys a region receives (one merge
    // of the cache per region at the end), and enough of them to spread the merge over every CU.
    const int want_pbits = std::max(0, std::min(PART_MAX_PBITS, log2_of((uint64_t)(rows >> 15) + 1)));
    if (h.pbits < want_pbits) VR_TRY(grow_table(h, std::max<uint64_t>(h.slots, 2048ull << want_pbits), s, want_pbits));
    const int P = 1 << h.pbits;
    const int E = merge_cache_entries(h);
    const int64_t budget = h.dev->lds_per_block - 1024;
    int K = 0, H = 0;
    for (int k = 2048; k >= 256; k -= 256) {
        const int hs = (int)next_pow2(2 * (uint64_t)k);
        if ((int64_t)scatter_lds_bytes(k, hs, S, P) <= budget) {
            K = k;
            H = hs;
            break;
        }
    }
    if ((int64_t)scatter_lds_bytes(K, H, S, P) > budget) return fail(VR_ERR_UNSUPPORTED, "state rows too wide for the scatter");
    const int grid_a = h.dev->num_cus;
    const int cap = h.dev->num_cus * 8;
    const uint64_t empty = EMPTY_KEY;
    a.t = view_of(h);
    PartArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.rl = L;
    pa.nprog = column_programs(h, L, pa.prog);
    pa.nparts = P;
    pa.nhot = K;
    pa.hot_slots = H;
    pa.grid_a = grid_a;
    pa.cache_e = E;

    VR_TRY(timer_begin(h, s));
    // the hot set, from a sample of the input
    if (K) {
        const uint64_t cand_n = 1ull << 20;
        VR_TRY(h.p_cand.ensure(h.dev, cand_n * sizeof(uint64_t)));
        VR_TRY(h.p_cand_count.ensure(h.dev, cand_n * sizeof(uint32_t)));
        VR_TRY(h.p_hist.ensure(h.dev, 264 * sizeof(uint32_t)));
        VR_TRY(h.p_hot_keys.ensure(h.dev, (size_t)K * sizeof(uint64_t)));
        VR_TRY(h.p_hot_part.ensure(h.dev, (size_t)grid_a * K * S * sizeof(uint64_t)));
        VR_TRY(h.p_hot_total.ensure(h.dev, (size_t)K * S * sizeof(uint64_t)));
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(cand_n, 256, cap)), dim3(256), 0, s, h.p_cand.as<uint64_t>(), cand_n, empty);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(cand_n / 2, 256, cap)), dim3(256), 0, s, h.p_cand_count.as<uint64_t>(),
                           cand_n / 2, 0ull);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(256), 0, s, h.p_hist.as<uint64_t>(), (uint64_t)132, 0ull);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(K, 256, cap)), dim3(256), 0, s, h.p_hot_keys.as<uint64_t>(),
                           (uint64_t)K, empty);
        const int64_t tiles = a.total_tiles - a.tile_lo;
        const int64_t nsamp = std::min<int64_t>(512, tiles);
        hipLaunchKernelGGL(hot_sample_kernel, dim3((unsigned)nsamp), dim3(HS_BLOCK), 0, s, a, nsamp, h.p_cand.as<uint64_t>(),
                           h.p_cand_count.as<uint32_t>(), cand_n - 1);
        hipLaunchKernelGGL(hot_hist_kernel, dim3(64), dim3(256), 0, s, h.p_cand_count.as<uint32_t>(), cand_n,
                           h.p_hist.as<uint32_t>());
        hipLaunchKernelGGL(hot_pick_kernel, dim3(64), dim3(256), 0, s, h.p_cand.as<uint64_t>(), h.p_cand_count.as<uint32_t>(),
                           cand_n, h.p_hist.as<uint32_t>(), K, h.p_hot_keys.as<uint64_t>(), h.p_hist.as<uint32_t>() + 256);
        VR_TRY_HIP(hipGetLastError());
        *launches += 7;
        pa.hot_keys = h.p_hot_keys.as<uint64_t>();
        pa.hot_part = h.p_hot_part.as<uint64_t>();
        pa.hot_total = h.p_hot_total.as<uint64_t>();
    }
    // the record pool and its ranges: the count pass, a scan down each region's column of counts,
    // a scan over the regions
    const uint64_t nblocks = pool_blocks(rows, (uint64_t)P);
    if (nblocks * PART_BLK >= (1ull << 32)) return fail(VR_ERR_UNSUPPORTED, "too many rows for one partitioned update");
    VR_TRY(h.p_pool.ensure(h.dev, nblocks * PART_BLK * L.words * sizeof(uint32_t)));
    // the count pass runs two workgroups per scatter workgroup when two fit on a CU
    const size_t lds_c = count_lds_bytes(K, H, P);
    pa.count_split = 2 * (int64_t)lds_c + 2048 <= h.dev->lds_per_block ? 2 : 1;
    VR_TRY(h.p_count.ensure(h.dev, (size_t)pa.count_split * grid_a * P * sizeof(uint32_t)));
    VR_TRY(h.p_base.ensure(h.dev, (size_t)grid_a * P * sizeof(uint32_t)));
    VR_TRY(h.p_rtot.ensure(h.dev, (size_t)P * sizeof(uint32_t)));
    VR_TRY(h.p_rstart.ensure(h.dev, (size_t)(P + 1) * sizeof(uint32_t)));
    VR_TRY(h.p_resume.ensure(h.dev, (size_t)P * sizeof(int32_t)));
    VR_TRY(h.p_pending.ensure(h.dev, 2 * nblocks * sizeof(uint64_t)));
    pa.pool = h.p_pool.as<uint32_t>();
    pa.count = h.p_count.as<uint32_t>();
    pa.base = h.p_base.as<uint32_t>();
    pa.rtot = h.p_rtot.as<uint32_t>();
    pa.rstart = h.p_rstart.as<uint32_t>();
    pa.resume_step = h.p_resume.as<int32_t>();
    pa.tile_batch = h.p_tile_batch.as<uint32_t>();
    pa.pending = h.p_pending.as<uint64_t>();
    pa.pool_blocks = nblocks;
    pa.pool_records = (uint32_t)(nblocks * PART_BLK);
    pa.u = a;
    const size_t lds_a = scatter_lds_bytes(K, H, S, P);
    const size_t lds_b = merge_lds_bytes(E, S);
    if (jit) {
        if (launch_jit(jit->count, (unsigned)(grid_a * pa.count_split), PA_BLOCK, lds_c, s, pa) != VR_OK) {
            // the module's kernels do not launch here: the generic kernels from the next call on
            (void)hipGetLastError();
            part_jit_disable();
            return fail(VR_ERR_DEVICE, "the partitioned update's compiled kernels failed to launch");
        }
    } else {
        VR_TRY(allow_lds(part_count_kernel<RuntimeProgs>, lds_c));
        hipLaunchKernelGGL(part_count_kernel<RuntimeProgs>, dim3(grid_a * pa.count_split), dim3(PA_BLOCK), lds_c, s, pa);
    }
    hipLaunchKernelGGL(part_col_scan_kernel, dim3((unsigned)((P + 63) / 64)), dim3(1024), 0, s, pa);
    hipLaunchKernelGGL(part_region_scan_kernel, dim3(1), dim3(1024), 0, s, pa);
    VR_TRY_HIP(hipGetLastError());
    *launches += 3;
    if (jit && launch_jit(jit->scatter, (unsigned)grid_a, PA_BLOCK, lds_a, s, pa) != VR_OK) {
        // the module's kernels do not launch here: the generic kernels from the next call on (this
        // call's tiles are cut for the compiled kernel's rows per thread)
        (void)hipGetLastError();
        part_jit_disable();
        return fail(VR_ERR_DEVICE, "the partitioned update's compiled kernels failed to launch");
    }
    *specialized = jit != nullptr;
    if (!jit) {
        switch (L.words) {
            case 4:
                VR_TRY(allow_lds(part_scatter_kernel<4, RuntimeProgs>, lds_a));
                hipLaunchKernelGGL((part_scatter_kernel<4, RuntimeProgs>), dim3(grid_a), dim3(PA_BLOCK), lds_a, s, pa);
                break;
            case 8:
                VR_TRY(allow_lds(part_scatter_kernel<8, RuntimeProgs>, lds_a));
                hipLaunchKernelGGL((part_scatter_kernel<8, RuntimeProgs>), dim3(grid_a), dim3(PA_BLOCK), lds_a, s, pa);
                break;
            case 12:
                VR_TRY(allow_lds(part_scatter_kernel<12, RuntimeProgs>, lds_a));
                hipLaunchKernelGGL((part_scatter_kernel<12, RuntimeProgs>), dim3(grid_a), dim3(PA_BLOCK), lds_a, s, pa);
                break;
            default:
                VR_TRY(allow_lds(part_scatter_kernel<16, RuntimeProgs>, lds_a));
                hipLaunchKernelGGL((part_scatter_kernel<16, RuntimeProgs>), dim3(grid_a), dim3(PA_BLOCK), lds_a, s, pa);
                break;
        }
    }
    if (K) hipLaunchKernelGGL(hot_reduce_kernel, dim3(grid_for((uint64_t)K * S, 256, cap)), dim3(256), 0, s, pa);
    VR_TRY_HIP(hipGetLastError());
    *launches += 3;
    // the merge: one workgroup per region; stopped regions resume after the table grows
    pa.resumed = 0;
    for (int round = 0;; round++) {
        pa.u.t = view_of(h);
        pa.region_limit = (uint32_t)((pa.u.t.rmask + 1) * 3 / 4);
        if (round) VR_TRY(timer_begin(h, s));
        if (jit) {
            VR_TRY(launch_jit(jit->merge, (unsigned)P, PB_BLOCK, lds_b, s, pa));
        } else {
            switch (L.words) {
                case 4:
                    VR_TRY(allow_lds(part_merge_kernel<4, RuntimeProgs>, lds_b));
                    hipLaunchKernelGGL((part_merge_kernel<4, RuntimeProgs>), dim3(P), dim3(PB_BLOCK), lds_b, s, pa);
                    break;
                case 8:
                    VR_TRY(allow_lds(part_merge_kernel<8, RuntimeProgs>, lds_b));
                    hipLaunchKernelGGL((part_merge_kernel<8, RuntimeProgs>), dim3(P), dim3(PB_BLOCK), lds_b, s, pa);
                    break;
                case 12:
                    VR_TRY(allow_lds(part_merge_kernel<12, RuntimeProgs>, lds_b));
                    hipLaunchKernelGGL((part_merge_kernel<12, RuntimeProgs>), dim3(P), dim3(PB_BLOCK), lds_b, s, pa);
                    break;
                default:
                    VR_TRY(allow_lds(part_merge_kernel<16, RuntimeProgs>, lds_b));
                    hipLaunchKernelGGL((part_merge_kernel<16, RuntimeProgs>), dim3(P), dim3(PB_BLOCK), lds_b, s, pa);
                    break;
            }
        }
        VR_TRY(timer_end(h, s));
        VR_TRY_HIP(hipGetLastError());
        (*launches)++;
        uint64_t c[C_WORDS];
        VR_TRY(read_counters(h, c, s));
        h.groups_upper = groups_of(c);
        if (!c[C_PSTOP]) break;
        if (round >= 40) return fail(VR_ERR_DEVICE, "the partitioned merge did not finish");
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h.counters.as<uint64_t>() + C_PSTOP, 1ull, 0ull);
        VR_TRY(grow_table(h, h.slots * 2, s, h.pbits));
        pa.resumed = 1;
    }
    return VR_OK;
}

// The partitioned update of nb device-resident batches (host descriptors: columns, selection, n).
static vr_status update_partitioned(AggImpl& h, BatchDesc* descs, int nb, int64_t rows, int32_t filter_col,
                                    int32_t filter_threshold, const RecLayout& L, const vr_launch_config* cfg,
                                    hipStream_t s) {
    UpdateArgs a = h.proto;
    a.filter_col = filter_col;
    a.filter_threshold = filter_threshold;
    // kernels compiled for this spec (the first call per spec compiles them: outside the timed part)
    const int32_t spec_mode = (cfg && cfg->specialize) ? cfg->specialize : h.cfg.specialize;
    const PartJit* jit = nullptr;
    if (spec_mode >= 0) {
        ColProg prog[VR_MAX_VALUES];
        const int np = column_programs(h, L, prog);
        jit = part_jit(h.dev->ordinal, prog, np, L, h.proto.ks.mode == KM_I64);
    }
    int rpt = (cfg && cfg->rows_per_thread) ? cfg->rows_per_thread : h.cfg.rows_per_thread;
    // the compiled kernels take PA_JIT_RPT rows per thread (a constant of theirs); the generic ones
    // the configured number, up to PA_MAX_RPT
    a.rows_per_thread = jit ? PA_JIT_RPT : std::max(1, std::min(PA_MAX_RPT, rpt));
    const int64_t T = (int64_t)PA_BLOCK * a.rows_per_thread;
    int64_t tiles = 0, seen = 0;
    for (int b = 0; b < nb; b++) {
        descs[b].row_base = h.rows_seen + seen;
        descs[b].tile_begin = tiles;
        tiles += (descs[b].n + T - 1) / T;
        seen += descs[b].n;
    }
    // descriptors and the tile -> batch map go up from page-locked buffers (descs may already be
    // the handle's own page-locked buffer: the batched entry point describes into it)
    const size_t dbytes = sizeof(BatchDesc) * (size_t)nb;
    if (descs != h.desc_host.as<BatchDesc>()) {
        VR_TRY(h.desc_host.acquire(dbytes));
        std::memcpy(h.desc_host.p, descs, dbytes);
    }
    VR_TRY(h.batch_descs.ensure(h.dev, dbytes));
    VR_TRY(h.desc_host.copy_to(h.batch_descs.p, dbytes, s));
    VR_TRY(h.tile_host.acquire(sizeof(uint32_t) * (size_t)std::max<int64_t>(tiles, 1)));
    uint32_t* tile_batch = h.tile_host.as<uint32_t>();
    for (int b = 0; b < nb; b++) {
        const int64_t t1 = b + 1 < nb ? descs[b + 1].tile_begin : tiles;
        for (int64_t t = descs[b].tile_begin; t < t1; t++) tile_batch[t] = (uint32_t)b;
    }
    VR_TRY(h.p_tile_batch.ensure(h.dev, sizeof(uint32_t) * (size_t)std::max<int64_t>(tiles, 1)));
    VR_TRY(h.tile_host.copy_to(h.p_tile_batch.p, sizeof(uint32_t) * (size_t)tiles, s));
    a.batches = h.batch_descs.as<BatchDesc>();
    a.num_batches = nb;
    a.one = descs[0];
    a.tile_lo = 0;
    a.total_tiles = tiles;
    a.epoch = ++h.epoch;
    int launches = 0;
    bool specialized = false;
    VR_TRY(run_partitioned(h, a, rows, L, s, &launches, jit, &specialized));
    if (h.has_f64_minmax) {  // over the final table
        a.t = view_of(h);
        hipLaunchKernelGGL(nan_payload_kernel<true>, dim3(grid_for(tiles, 1, h.dev->num_cus * 4)), dim3(PA_BLOCK), 0, s, a);
        VR_TRY_HIP(hipGetLastError());
        launches++;
    }
    VR_TRY_HIP(hipStreamSynchronize(s));  // the descriptors are rewritten by the next call
    h.dev->count(&vr_stats::partitioned_updates, 1);
    if (specialized) h.dev->count(&vr_stats::specialized_updates, 1);
    h.rows_seen += rows;
    h.rows_since_reset += (uint64_t)rows;
    h.order_count = -1;
    h.dev->count(&vr_stats::kernel_launches, launches);
    return VR_OK;
}

static vr_status run_update(AggImpl& h, BatchDesc& b, int64_t n, hipStream_t s, const vr_launch_config* cfg) {
    RecLayout layout;
    if (partitioned_applies(h, n, &layout, cfg))
        return update_partitioned(h, &b, 1, n, h.proto.filter_col, h.proto.filter_threshold, layout, cfg, s);
    UpdateArgs a = h.proto;
    b.row_base = h.rows_seen;
    b.tile_begin = 0;
    a.one = b;
    a.num_batches = 1;
    int grid, block;
    size_t lds_bytes;
    launch_shape(h, cfg, n, a, &grid, &block, &lds_bytes, false);
    const int64_t T = (int64_t)block * a.rows_per_thread;
    a.tile_lo = 0;
    a.total_tiles = (n + T - 1) / T;
    a.epoch = ++h.epoch;
    int launches = 0;
    if (h.ordered || n <= GUARD_ROWS) {
        // an ordered fold needs each row's slot: one launch, no growth inside it
        VR_TRY(ensure_capacity(h, n, s));
        if (h.ordered) {
            VR_TRY(h.row_slot.ensure(h.dev, (size_t)n * sizeof(uint32_t)));
            a.row_slot = h.row_slot.as<uint32_t>();
        }
        a.t = view_of(h);
        VR_TRY(timer_begin(h, s));
        hipLaunchKernelGGL(agg_update_kernel<false>, dim3(grid), dim3(block), lds_bytes, s, a);
        VR_TRY(timer_end(h, s));
        VR_TRY_HIP(hipGetLastError());
        launches++;
        h.groups_upper += n;
    } else {
        VR_TRY(launch_guarded<false>(h, a, grid, block, lds_bytes, s, &launches));
    }
    if (h.ordered && n > 0) {
        // order the batch's rows by slot (stable: row order kept inside a slot), then fold
        VR_TRY(h.sort_keys.ensure(h.dev, (size_t)n * sizeof(uint32_t)));
        VR_TRY(h.sort_vals.ensure(h.dev, (size_t)n * sizeof(uint32_t)));
        VR_TRY(h.sort_vals_in.ensure(h.dev, (size_t)n * sizeof(uint32_t)));
        size_t tmp = 0;
        VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, a.row_slot, h.sort_keys.as<uint32_t>(),
                                                      h.sort_vals_in.as<uint32_t>(), h.sort_vals.as<uint32_t>(),
                                                      (int)n, 0, 32, s));
        VR_TRY(h.cub_tmp.ensure(h.dev, tmp));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(iota_u32_kernel), dim3(grid_for(n, 256, h.dev->num_cus * 16)), dim3(256), 0, s,
                           h.sort_vals_in.as<uint32_t>(), n);
        VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(h.cub_tmp.p, tmp, a.row_slot, h.sort_keys.as<uint32_t>(),
                                                      h.sort_vals_in.as<uint32_t>(), h.sort_vals.as<uint32_t>(),
                                                      (int)n, 0, 32, s));
        hipLaunchKernelGGL(ordered_fold_kernel, dim3(grid_for(n, 256, h.dev->num_cus * 16)), dim3(256), 0, s,
                           h.sort_keys.as<uint32_t>(), h.sort_vals.as<uint32_t>(), n, a);
        VR_TRY_HIP(hipGetLastError());
        launches += 3;
    }
    if (h.has_f64_minmax) {  // over the final table: after any growth
        a.t = view_of(h);
        hipLaunchKernelGGL(nan_payload_kernel<false>, dim3(grid_for(n, 256, h.dev->num_cus * 8)), dim3(256), 0, s, a);
        VR_TRY_HIP(hipGetLastError());
        launches++;
    }
    h.rows_seen += n;
    h.rows_since_reset += (uint64_t)n;
    h.order_count = -1;
    h.dev->count(&vr_stats::kernel_launches, launches);
    return VR_OK;
}

// The emission order: occupied slots in slot order (then the reserved rows), or by first-seen row.
static vr_status build_order(AggImpl& h, hipStream_t s) {
    if (h.order_count >= 0) return VR_OK;
    uint64_t c[C_WORDS];
    VR_TRY(read_counters(h, c, s));
    int64_t g = groups_of(c);
    VR_TRY(h.order.ensure(h.dev, (size_t)(g + 2) * sizeof(uint64_t)));
    const uint64_t nblocks = (h.slots + ORDER_CHUNK - 1) / ORDER_CHUNK;
    VR_TRY(h.order_tmp.ensure(h.dev, (size_t)(2 * nblocks + 2) * sizeof(uint64_t)));
    uint64_t* counts = h.order_tmp.as<uint64_t>();
    uint64_t* offsets = counts + nblocks + 1;
    const int W = h.proto.ks.width;
    hipLaunchKernelGGL(count_occupied_kernel, dim3((unsigned)nblocks), dim3(256), 0, s, h.cells.as<uint64_t>(), W,
                       h.slots, counts);
    size_t tmp = 0;
    VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, counts, offsets, (int)nblocks, s));
    VR_TRY(h.cub_tmp.ensure(h.dev, tmp));
    VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(h.cub_tmp.p, tmp, counts, offsets, (int)nblocks, s));
    hipLaunchKernelGGL(write_occupied_kernel, dim3((unsigned)nblocks), dim3(256), 0, s, h.cells.as<uint64_t>(), W,
                       h.slots, offsets, h.order.as<uint64_t>(), (uint64_t)(g + 2));
    // the occupied slots are the groups counted at insertion; the order's own count is checked on
    // the device (a difference fails the emission's closing read of the counters), with no wait here
    const uint64_t selected = c[C_GROUPS];
    uint64_t extra[2] = {0, 0};
    int nx = 0;
    if (c[C_RESERVED] & RES_SENTINEL) extra[nx++] = h.slots;
    if (c[C_RESERVED] & RES_NULL) extra[nx++] = h.slots + 1;
    hipLaunchKernelGGL(order_tail_kernel, dim3(1), dim3(64), 0, s, counts, offsets, nblocks, selected,
                       h.order.as<uint64_t>(), extra[0], extra[1], nx, h.counters.as<uint64_t>());
    VR_TRY_HIP(hipGetLastError());
    const int64_t m = (int64_t)selected + nx;  // == g
    if (h.det && m > 1) {
        int off = -1;
        for (int k = 0; k < h.proto.nops; k++)
            if (h.proto.ops[k].code == OP_FIRST_ROW) off = h.proto.ops[k].off;
        VR_TRY(h.firsts.ensure(h.dev, (size_t)m * sizeof(uint64_t)));
        VR_TRY(h.firsts_sorted.ensure(h.dev, (size_t)m * sizeof(uint64_t)));
        VR_TRY(h.order_tmp.ensure(h.dev, (size_t)m * sizeof(uint64_t)));
        hipLaunchKernelGGL(first_rows_kernel, dim3(grid_for(m, 256, h.dev->num_cus * 16)), dim3(256), 0, s,
                           h.order.as<uint64_t>(), m, h.state.as<uint64_t>(), h.stride, off, h.firsts.as<uint64_t>());
        size_t t2 = 0;
        VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, h.firsts.as<uint64_t>(), h.firsts_sorted.as<uint64_t>(),
                                                      h.order.as<uint64_t>(), h.order_tmp.as<uint64_t>(), (int)m, 0, 64, s));
        VR_TRY(h.cub_tmp.ensure(h.dev, t2));
        VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(h.cub_tmp.p, t2, h.firsts.as<uint64_t>(), h.firsts_sorted.as<uint64_t>(),
                                                      h.order.as<uint64_t>(), h.order_tmp.as<uint64_t>(), (int)m, 0, 64, s));
        std::swap(h.order.p, h.order_tmp.p);
        std::swap(h.order.bytes, h.order_tmp.bytes);
    }
    // no wait: the emission's kernel is queued after these on the same stream, and the emission
    // ends with a synchronous read of the counters
    h.order_count = m;
    return VR_OK;
}

static vr_status enter(AggImpl& h, hipStream_t s) {
    if (h.last_stream && h.last_stream != s) VR_TRY_HIP(hipStreamWaitEvent(s, h.last_ev, 0));
    return VR_OK;
}

static vr_status leave(AggImpl& h, hipStream_t s) {
    if (!h.last_ev) VR_TRY_HIP(hipEventCreateWithFlags(&h.last_ev, hipEventDisableTiming));
    VR_TRY_HIP(hipEventRecord(h.last_ev, s));
    h.last_stream = s;
    return VR_OK;
}

}  // namespace vr

using namespace vr;

struct vr_agg {
    DeviceRef owner;  // destroyed after the state below
    AggImpl h;
    std::vector<vr_agg_graph*> graphs;  // live graphs captured from this handle (detached at destroy)
};

// A captured batched update: a HIP graph of one whole step over a fixed batch set (the table
// cleared, every batch folded with one unguarded persistent launch, the NaN-payload pass), with the
// device copies of the batch descriptors and key dictionary maps it reads.
struct vr_agg_graph {
    DeviceRef owner;     // destroyed after the buffers below
    vr_agg* agg = nullptr;  // null once the aggregate is destroyed: launches are then refused
    hipEvent_t done = nullptr;  // recorded after each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf descs, remaps;
    int64_t rows = 0;
    // the table the graph was captured against: a replay after the table moved is refused
    void *cells = nullptr, *state = nullptr, *counters = nullptr;
    uint64_t slots = 0;
    ~vr_agg_graph() {
        if (done) {
            (void)hipEventSynchronize(done);  // a replay may still read the graph's buffers
            (void)hipEventDestroy(done);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

extern "C" {

vr_status vr_agg_create(vr_device* dev, const vr_agg_spec* spec, vr_agg** out) {
    if (!dev || !spec || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (spec->num_keys > 0 && !spec->key_types) return fail(VR_ERR_INVALID_ARGUMENT, "null key types");
    std::unique_ptr<vr_agg> a(new vr_agg());
    a->owner.hold(dev);
    AggImpl& h = a->h;
    h.dev = &dev->d;
    VR_TRY_HIP(hipSetDevice(h.dev->ordinal));
    if (spec->num_keys < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative key count");
    h.ungrouped = spec->num_keys == 0;
    if (h.ungrouped) {  // one group: a constant key the library supplies
        h.key_types.assign(1, VR_INT64);
        h.key_nullable.assign(1, 0);
    } else {
        h.key_types.assign(spec->key_types, spec->key_types + spec->num_keys);
        h.key_nullable.assign(spec->num_keys, 1);
        if (spec->key_nullable)
            for (int c = 0; c < spec->num_keys; c++) h.key_nullable[c] = spec->key_nullable[c] ? 1 : 0;
    }
    if (spec->num_values > 0) h.value_types.assign(spec->value_types, spec->value_types + spec->num_values);
    if (spec->num_aggs > 0) h.aggs.assign(spec->aggs, spec->aggs + spec->num_aggs);
    h.det = spec->deterministic != 0;
    h.ordered = h.det;
    if (spec->max_load_percent < 0 || spec->max_load_percent > 90)
        return fail(VR_ERR_INVALID_ARGUMENT, "max_load_percent outside [0, 90]");
    if (spec->max_load_percent) h.max_load = spec->max_load_percent / 100.0;
    VR_TRY(compile_spec(h));
    for (int k = 0; k < h.proto.nops; k++)
        if (h.proto.ops[k].code == OP_SUM_I64 || h.proto.ops[k].code == OP_SUM_D128) h.half_sums = true;
    h.strs.resize(h.key_types.size());
    VR_TRY(init_table(h, h.ungrouped ? 1 : spec->capacity_hint));
    if (h.ungrouped) {
        VR_TRY(insert_single_group(h, h.dev->stream));
        VR_TRY_HIP(hipStreamSynchronize(h.dev->stream));
    }
    *out = a.release();
    return VR_OK;
}

// The bound of the 32-bit-piece sums: `rows` more rows must keep every group below 2^32 rows.
static vr_status check_row_bound(const AggImpl& h, uint64_t rows) {
    if (h.half_sums && h.rows_since_reset + rows >= (1ull << 32))
        return fail(VR_ERR_UNSUPPORTED, "an aggregate with integer or decimal sums takes fewer than 2^32 rows between "
                                        "resets (emit and reset the handle, or run the CPU kernels)");
    return VR_OK;
}

vr_status vr_agg_update(vr_agg* agg, const vr_column* keys, const vr_column* values, int64_t n,
                        const uint8_t* selection, void* stream) {
    if (!agg || (!keys && !agg->h.ungrouped)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative row count");
    AggImpl& h = agg->h;
    if (n == 0) return VR_OK;
    if ((uint64_t)n >= (1ull << 32) && h.ordered)
        return fail(VR_ERR_UNSUPPORTED, "deterministic batches and batches of a 128-bit MIN/MAX are below 2^32 rows");
    VR_TRY(check_row_bound(h, (uint64_t)n));
    hipStream_t s = h.dev->pick(stream);
    VR_TRY(enter(h, s));
    vr_column zk;
    if (h.ungrouped) {
        VR_TRY(zero_key_column(h, n, s, &zk));
        keys = &zk;
    }
    BatchDesc b;
    StagePlan plan;
    KeyIds ids(s);
    int32_t mem;
    VR_TRY(describe(h, keys, values, n, selection, b, plan, ids.bufs, &mem, s));
    int slot;
    VR_TRY(stage(h, plan, s, &slot));
    VR_TRY(run_update(h, b, n, s, nullptr));
    if (slot >= 0) VR_TRY(h.staging.done(slot, s));
    VR_TRY(leave(h, s));
    h.dev->count(&vr_stats::calls, 1);
    h.dev->count(&vr_stats::input_rows, n);
    return VR_OK;
}

vr_status vr_agg_update_filtered(vr_agg* agg, const vr_column* keys, const vr_column* values, int64_t n,
                                 int32_t filter_col, int32_t filter_threshold, void* stream) {
    if (!agg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    if (filter_col >= (int)h.value_types.size() || (filter_col >= 0 && h.value_types[filter_col] != VR_INT32))
        return fail(VR_ERR_INVALID_ARGUMENT, "the filter column must be an INT32 value column");
    h.proto.filter_col = filter_col;
    h.proto.filter_threshold = filter_threshold;
    vr_status st = vr_agg_update(agg, keys, values, n, nullptr, stream);
    h.proto.filter_col = -1;
    return st;
}

// Host-resident batch sets: consecutive batches are staged in chunks of up to 256 MiB (page-locked
// buffers copied by DMA as they are, pageable ones through the pinned buffer by several host
// threads, consecutive pieces in one copy) on the copy stream, and each chunk is folded by the
// device-resident batched update while the copies of the next chunk run.
static vr_status update_batches_from_host(vr_agg* agg, const vr_batch_set* bs, const vr_launch_config* cfg, hipStream_t s) {
    AggImpl& h = agg->h;
    const int nb = bs->num_batches;
    const int nk = (int)h.key_types.size(), nv = (int)h.value_types.size();
    constexpr size_t CHUNK = (size_t)256 << 20;
    // the bytes of a host column of n rows to stage, and its pieces (pointers patched in `col`)
    auto col_bytes = [](const vr_column& col, int64_t n) -> size_t {
        if (col.mem != VR_MEM_HOST || n <= 0) return 0;
        size_t b = col.validity ? bitmap_bytes(n) + 256 : 0;
        if (col.type == VR_UTF8 && !col.dictionary) return b + (size_t)(n + 1) * 4 + (col.offsets ? (size_t)std::max(col.offsets[n], 0) : 0) + 512;
        return b + lane_bytes(col.type, n) + 256;
    };
    auto col_pieces = [](vr_column& col, int64_t n, std::vector<StagePiece>& pieces) {
        if (col.mem != VR_MEM_HOST || n <= 0) return;
        col.mem = VR_MEM_DEVICE;
        if (col.validity) pieces.push_back({col.validity, bitmap_bytes(n), (const void**)&col.validity});
        if (col.type == VR_UTF8 && !col.dictionary) {
            if (!col.offsets) return;  // refused by the update
            pieces.push_back({col.offsets, (size_t)(n + 1) * 4, (const void**)&col.offsets});
            const size_t nbytes = (size_t)std::max(col.offsets[n], 0);
            if (nbytes) pieces.push_back({col.data, nbytes, &col.data});
        } else if (lane_bytes(col.type, n)) {
            pieces.push_back({col.data, lane_bytes(col.type, n), &col.data});
        }
    };
    struct Chunk {
        int b0 = 0, b1 = 0;
        std::vector<vr_column> keys, vals;
        std::vector<StagePiece> pieces;
        int slot = -1;
    };
    std::vector<std::pair<int, int>> bounds;
    for (int b0 = 0; b0 < nb;) {
        size_t bytes = 0;
        int b1 = b0;
        while (b1 < nb) {
            size_t bb = 0;
            for (int c = 0; c < nk; c++) bb += col_bytes(bs->keys[(size_t)b1 * nk + c], bs->lengths[b1]);
            for (int c = 0; c < nv; c++) bb += col_bytes(bs->values[(size_t)b1 * nv + c], bs->lengths[b1]);
            if (b1 > b0 && bytes + bb > CHUNK) break;
            bytes += bb;
            b1++;
        }
        bounds.push_back({b0, b1});
        b0 = b1;
    }
    auto prepare = [&](Chunk& ch, int k) -> vr_status {
        ch.b0 = bounds[(size_t)k].first;
        ch.b1 = bounds[(size_t)k].second;
        const int m = ch.b1 - ch.b0;
        ch.keys.assign(bs->keys + (size_t)ch.b0 * nk, bs->keys + (size_t)ch.b1 * nk);
        ch.vals.assign(bs->values + (size_t)ch.b0 * nv, bs->values + (size_t)ch.b1 * nv);
        ch.pieces.clear();
        for (int b = 0; b < m; b++) {
            const int64_t n = bs->lengths[ch.b0 + b];
            for (int c = 0; c < nk; c++) col_pieces(ch.keys[(size_t)b * nk + c], n, ch.pieces);
            for (int c = 0; c < nv; c++) col_pieces(ch.vals[(size_t)b * nv + c], n, ch.pieces);
        }
        return h.staging.stage(h.dev, ch.pieces, s, &ch.slot, true);
    };
    Chunk cur, next;
    VR_TRY(prepare(cur, 0));
    for (size_t k = 0; k < bounds.size(); k++) {
        if (cur.slot >= 0) VR_TRY(h.staging.ready(cur.slot, s));
        // the next chunk's host copies and DMA on another thread, beside this chunk's fold
        std::future<vr_status> staged;
        if (k + 1 < bounds.size())
            staged = std::async(std::launch::async, [&, k]() -> vr_status {
                if (hipSetDevice(h.dev->ordinal) != hipSuccess) return VR_ERR_DEVICE;
                return prepare(next, (int)k + 1);
            });
        vr_batch_set cb = *bs;
        cb.num_batches = cur.b1 - cur.b0;
        cb.keys = cur.keys.data();
        cb.values = cur.vals.data();
        cb.lengths = bs->lengths + cur.b0;
        const vr_status st = vr_agg_update_batches(agg, &cb, cfg, s);
        const vr_status st_next = staged.valid() ? staged.get() : VR_OK;
        VR_TRY(st);
        if (st_next != VR_OK) return fail(st_next, "staging a chunk of host-resident batches failed");
        if (cur.slot >= 0) VR_TRY(h.staging.done(cur.slot, s));
        std::swap(cur, next);
    }
    // every copy out of the caller's page-locked buffers has ended before the call returns
    VR_TRY_HIP(hipStreamSynchronize(h.dev->copy_stream));
    return VR_OK;
}

vr_status vr_agg_update_batches(vr_agg* agg, const vr_batch_set* bs, const vr_launch_config* cfg, void* stream) {
    if (!agg || !bs) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    const int nb = bs->num_batches;
    if (nb <= 0) return VR_OK;
    const int nk = (int)h.key_types.size(), nv = (int)h.value_types.size();
    if (bs->filter_col >= nv || (bs->filter_col >= 0 && h.value_types[bs->filter_col] != VR_INT32))
        return fail(VR_ERR_INVALID_ARGUMENT, "the filter column must be an INT32 value column");
    uint64_t all_rows = 0;
    for (int b = 0; b < nb; b++) {
        if (bs->lengths[b] < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative batch length");
        all_rows += (uint64_t)bs->lengths[b];
    }
    VR_TRY(check_row_bound(h, all_rows));
    hipStream_t s = h.dev->pick(stream);
    vr_batch_set with_keys;
    std::vector<vr_column> zkeys;
    if (h.ungrouped) {  // the constant key column of every batch
        int64_t longest = 0;
        for (int b = 0; b < nb; b++) longest = std::max(longest, bs->lengths[b]);
        vr_column zk;
        VR_TRY(zero_key_column(h, longest, s, &zk));
        zkeys.assign(nb, zk);
        for (int b = 0; b < nb; b++) zkeys[b].length = bs->lengths[b];
        with_keys = *bs;
        with_keys.keys = zkeys.data();
        bs = &with_keys;
    }
    // host-resident columns: staged in chunks, each folded as a device-resident set
    bool host = false;
    for (int b = 0; b < nb && !host; b++) {
        for (int c = 0; c < nk; c++) host |= bs->keys[(size_t)b * nk + c].mem == VR_MEM_HOST;
        for (int c = 0; c < nv; c++) host |= bs->values[(size_t)b * nv + c].mem == VR_MEM_HOST;
    }
    if (host && !h.ordered) return update_batches_from_host(agg, bs, cfg, s);
    // One persistent launch needs device-resident columns and one dictionary per UTF8 key; anything
    // else (and an ordered fold, which orders each batch) takes the per-batch path.
    bool fused = !h.ordered;
    for (int b = 0; b < nb && fused; b++) {
        for (int c = 0; c < nk; c++) {
            const vr_column& k = bs->keys[(size_t)b * nk + c];
            if (k.mem != VR_MEM_DEVICE) fused = false;
            if (k.type == VR_UTF8 && (!k.dictionary || !bs->keys[c].dictionary ||
                                      k.dictionary->data != bs->keys[c].dictionary->data))
                fused = false;
        }
        for (int c = 0; c < nv; c++)
            if (bs->values[(size_t)b * nv + c].mem != VR_MEM_DEVICE) fused = false;
    }
    if (!fused) {
        for (int b = 0; b < nb; b++)
            VR_TRY(vr_agg_update_filtered(agg, bs->keys + (size_t)b * nk, bs->values + (size_t)b * nv,
                                          bs->lengths[b], bs->filter_col, bs->filter_threshold, stream));
        return VR_OK;
    }
    int64_t total = 0;
    for (int b = 0; b < nb; b++) {
        if (bs->lengths[b] < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative batch length");
        total += bs->lengths[b];
    }
    RecLayout layout;
    if (partitioned_applies(h, total, &layout, cfg)) {
        VR_TRY(enter(h, s));
        VR_TRY(h.desc_host.acquire(sizeof(BatchDesc) * (size_t)nb));  // described in place, copied by DMA
        BatchDesc* pdescs = h.desc_host.as<BatchDesc>();
        KeyIds ids(s);
        for (int b = 0; b < nb; b++) {
            StagePlan plan;
            int32_t mem;
            VR_TRY(describe(h, bs->keys + (size_t)b * nk, bs->values + (size_t)b * nv, bs->lengths[b], nullptr, pdescs[b],
                            plan, ids.bufs, &mem, s));
        }
        VR_TRY(update_partitioned(h, pdescs, nb, total, bs->filter_col, bs->filter_threshold, layout, cfg, s));
        h.dev->count(&vr_stats::calls, 1);
        h.dev->count(&vr_stats::input_rows, total);
        return leave(h, s);
    }
    UpdateArgs a = h.proto;
    a.filter_col = bs->filter_col;
    a.filter_threshold = bs->filter_threshold;
    // the tile size first: batches are cut into tiles of block * rows_per_thread rows
    int grid, block;
    size_t lds_bytes;
    a.total_tiles = 1;
    launch_shape(h, cfg, 0, a, &grid, &block, &lds_bytes, true);
    const int64_t T = (int64_t)block * a.rows_per_thread;
    VR_TRY(enter(h, s));
    VR_TRY_HIP(hipStreamSynchronize(s));  // the previous batched launch may still read the descriptors
    static thread_local std::vector<BatchDesc> descs;
    descs.resize(nb);
    int64_t tiles = 0, rows = 0;
    KeyIds ids(s);
    for (int b = 0; b < nb; b++) {
        StagePlan plan;
        int32_t mem;
        int64_t n = bs->lengths[b];
        if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative batch length");
        VR_TRY(describe(h, bs->keys + (size_t)b * nk, bs->values + (size_t)b * nv, n, nullptr, descs[b], plan, ids.bufs,
                        &mem, s));
        descs[b].row_base = h.rows_seen + rows;
        descs[b].tile_begin = tiles;
        tiles += (n + T - 1) / T;
        rows += n;
    }
    if (rows == 0) return VR_OK;
    VR_TRY(h.batch_descs.ensure(h.dev, sizeof(BatchDesc) * (size_t)nb));
    VR_TRY_HIP(hipMemcpyAsync(h.batch_descs.p, descs.data(), sizeof(BatchDesc) * (size_t)nb, hipMemcpyHostToDevice, s));
    // One persistent launch over every tile; a long input takes the guarded launch.
    int launches = 0;
    const int cap = batches_cap(h, cfg);
    grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    a.batches = h.batch_descs.as<BatchDesc>();
    a.num_batches = nb;
    a.tile_lo = 0;
    a.total_tiles = tiles;
    a.epoch = ++h.epoch;
    if (rows <= GUARD_ROWS) {
        VR_TRY(ensure_capacity(h, rows, s));
        a.t = view_of(h);
        VR_TRY(timer_begin(h, s));
        hipLaunchKernelGGL(agg_update_kernel<true>, dim3(grid), dim3(block), lds_bytes, s, a);
        VR_TRY(timer_end(h, s));
        VR_TRY_HIP(hipGetLastError());
        launches++;
        h.groups_upper += rows;
    } else {
        VR_TRY(launch_guarded<true>(h, a, grid, block, lds_bytes, s, &launches));
    }
    if (h.has_f64_minmax) {  // over the final table: after any growth
        a.t = view_of(h);
        hipLaunchKernelGGL(nan_payload_kernel<true>, dim3(grid), dim3(block), 0, s, a);
        VR_TRY_HIP(hipGetLastError());
        launches++;
    }
    h.rows_seen += rows;
    h.rows_since_reset += (uint64_t)rows;
    h.order_count = -1;
    h.dev->count(&vr_stats::kernel_launches, launches);
    h.dev->count(&vr_stats::calls, 1);
    h.dev->count(&vr_stats::input_rows, rows);
    return leave(h, s);
}

vr_status vr_agg_set_launch_config(vr_agg* agg, const vr_launch_config* cfg) {
    if (!agg || !cfg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    vr_launch_config& c = agg->h.cfg;
    if (cfg->rows_per_thread < 0 || cfg->rows_per_thread > 64 || cfg->blocks_per_cu < 0 || cfg->blocks_per_cu > 32 ||
        cfg->lds_entries < 0 || cfg->lds_entries > 2048 || cfg->flush_tiles < 0 || cfg->partitioned < -1 ||
        cfg->partitioned > 1 || cfg->specialize < -1 || cfg->specialize > 1)
        return fail(VR_ERR_INVALID_ARGUMENT, "launch config out of range");
    if (cfg->rows_per_thread) c.rows_per_thread = cfg->rows_per_thread;
    if (cfg->blocks_per_cu) c.blocks_per_cu = cfg->blocks_per_cu;
    if (cfg->lds_preaggregate >= 0) c.lds_preaggregate = cfg->lds_preaggregate;
    if (cfg->lds_entries) c.lds_entries = cfg->lds_entries;
    if (cfg->flush_tiles) c.flush_tiles = cfg->flush_tiles;
    if (cfg->partitioned) c.partitioned = cfg->partitioned;
    if (cfg->specialize) c.specialize = cfg->specialize;
    return VR_OK;
}

vr_status vr_agg_num_groups(vr_agg* agg, int64_t* groups) {
    if (!agg || !groups) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t c[C_WORDS];
    VR_TRY(enter(agg->h, agg->h.dev->stream));
    VR_TRY(read_counters(agg->h, c, agg->h.dev->stream));
    VR_TRY(leave(agg->h, agg->h.dev->stream));
    *groups = groups_of(c);
    return VR_OK;
}

vr_status vr_agg_output_layout(vr_agg* agg, int32_t* num_columns, int32_t* types, int32_t max_columns) {
    if (!agg || !num_columns) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    *num_columns = (int32_t)agg->h.out_types.size();
    for (int i = 0; types && i < max_columns && i < (int)agg->h.out_types.size(); i++) types[i] = agg->h.out_types[i];
    return VR_OK;
}

vr_status vr_agg_emit(vr_agg* agg, int64_t from, int64_t to, vr_out_column* keys, vr_out_column* out, void* stream) {
    if (!agg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    hipStream_t s = h.dev->pick(stream);
    VR_TRY(enter(h, s));
    VR_TRY(build_order(h, s));
    if (from < 0 || to < from || to > h.order_count) {
        (void)leave(h, s);  // the order's kernels are queued: later calls on other streams wait for them
        return fail(VR_ERR_INVALID_ARGUMENT, "emission range outside the groups");
    }
    const int64_t m = to - from;
    if (m == 0) return leave(h, s);
    const int nk = h.ungrouped ? 0 : (int)h.key_types.size(), no = (int)h.out_types.size();
    EmitArgs a{};
    a.t = view_of(h);
    a.ks = h.proto.ks;
    if (h.ungrouped) a.ks.nkeys = 0;  // no key columns out
    a.order = h.order.as<uint64_t>();
    a.from = from;
    a.m = m;
    a.naggs = (int)h.emits.size();
    for (int k = 0; k < a.naggs; k++) a.aggs[k] = h.emits[k];
    a.flags_off = h.proto.flags_off;
    // outputs: caller's device buffers directly, host ones through a device staging area
    struct Back {
        void* host;
        size_t off, bytes;
    };
    std::vector<Back> back;
    size_t need = 0;
    auto plan = [&](const vr_out_column& oc, int32_t type, OutCol& d) -> vr_status {
        if (oc.capacity < m) return fail(VR_ERR_CAPACITY, "output column holds fewer rows than emitted");
        d.type = type;
        size_t vb = bitmap_bytes(m), db = lane_bytes(type == VR_UTF8 ? VR_INT32 : type, m);
        if (oc.mem == VR_MEM_DEVICE) {
            d.valid = oc.validity;
            d.data = oc.data;
        } else {
            if (oc.validity) {
                need = align_up(need, 256);
                back.push_back({oc.validity, need, vb});
                need += vb;
            }
            need = align_up(need, 256);
            back.push_back({oc.data, need, db});
            need += db;
        }
        return VR_OK;
    };
    for (int c = 0; c < nk; c++) VR_TRY(plan(keys[c], h.key_types[c], a.keys[c]));
    for (int c = 0; c < no; c++) VR_TRY(plan(out[c], h.out_types[c], a.outs[c]));
    if (need) {
        VR_TRY(h.out_stage.ensure(h.dev, need));
        uint8_t* base = h.out_stage.as<uint8_t>();
        size_t k = 0;
        auto bind = [&](const vr_out_column& oc, OutCol& d) {
            if (oc.mem == VR_MEM_DEVICE) return;
            if (oc.validity) d.valid = base + back[k++].off;
            else d.valid = nullptr;
            d.data = base + back[k++].off;
        };
        for (int c = 0; c < nk; c++) bind(keys[c], a.keys[c]);
        for (int c = 0; c < no; c++) bind(out[c], a.outs[c]);
    }
    const unsigned eg = (unsigned)((m + EMIT_BLOCK - 1) / EMIT_BLOCK);
    if (emit_pitch(h.stride) <= EMIT_STAGE_MAX_PITCH)
        hipLaunchKernelGGL(agg_emit_kernel<true>, dim3(eg), dim3(EMIT_BLOCK),
                           (size_t)EMIT_BLOCK * emit_pitch(h.stride) * sizeof(uint64_t), s, a);
    else
        hipLaunchKernelGGL(agg_emit_kernel<false>, dim3(eg), dim3(EMIT_BLOCK), 0, s, a);
    VR_TRY_HIP(hipGetLastError());
    for (auto& bk : back) {
        VR_TRY_HIP(hipMemcpyAsync(bk.host, h.out_stage.as<uint8_t>() + bk.off, bk.bytes, hipMemcpyDeviceToHost, s));
        h.dev->count(&vr_stats::d2h_bytes, (int64_t)bk.bytes);
    }
    uint64_t c[C_WORDS];
    VR_TRY(read_counters(h, c, s));
    VR_TRY(leave(h, s));
    h.dev->count(&vr_stats::output_rows, m);
    if (c[C_ERROR]) {
        uint64_t e = c[C_ERROR];
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h.counters.as<uint64_t>() + C_ERROR, 1ull, 0ull);
        VR_TRY_HIP(hipStreamSynchronize(s));
        if (e & ERR_LONG_OVERFLOW) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "long overflow");
        if (e & ERR_AVG_OF_OVERFLOW) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "decimal average over an overflowed sum");
        if (e & ERR_SUM_OF_OVERFLOW) return fail(VR_ERR_ARITHMETIC_OVERFLOW, "overflow in sum of decimals");
        if (e & ERR_WIDE_TOTAL)
            return fail(VR_ERR_UNSUPPORTED, "ungrouped decimal average over a total beyond 128 bits runs on the CPU kernels");
        return fail(VR_ERR_DECIMAL_OVERFLOW, "decimal value past its precision");
    }
    return VR_OK;
}

vr_status vr_agg_key_dictionary(vr_agg* agg, int32_t c, int32_t* offsets, int64_t offsets_capacity, uint8_t* bytes,
                                int64_t bytes_capacity, int64_t* count, int64_t* num_bytes) {
    if (!agg || !count || !num_bytes) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    if (c < 0 || c >= (int)h.key_types.size() || h.key_types[c] != VR_UTF8)
        return fail(VR_ERR_INVALID_ARGUMENT, "not a UTF8 key column");
    std::vector<int32_t> offs;
    std::vector<uint8_t> data;
    VR_TRY(h.strs[c].table.read(&offs, &data, h.dev->stream));
    *count = (int64_t)offs.size() - 1;
    *num_bytes = (int64_t)data.size();
    if (!offsets || !bytes) return VR_OK;
    if (offsets_capacity < *count + 1 || bytes_capacity < *num_bytes) return fail(VR_ERR_CAPACITY, "dictionary buffers too small");
    std::memcpy(offsets, offs.data(), offs.size() * sizeof(int32_t));
    if (!data.empty()) std::memcpy(bytes, data.data(), data.size());
    return VR_OK;
}

vr_status vr_agg_prewarm(vr_device* dev, const vr_agg_spec* spec) {
    if (!dev || !spec) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    vr_agg* agg = nullptr;
    VR_TRY(vr_agg_create(dev, spec, &agg));
    std::unique_ptr<vr_agg, vr_status (*)(vr_agg*)> guard(agg, vr_agg_destroy);
    AggImpl& h = agg->h;
    RecLayout L;
    // an input large enough for the partitioned update; the spec decides whether it applies
    if (h.cfg.specialize < 0 || !partitioned_applies(h, h.part_min_rows, &L, nullptr)) return VR_OK;
    ColProg prog[VR_MAX_VALUES];
    const int np = column_programs(h, L, prog);
    (void)part_jit(h.dev->ordinal, prog, np, L, h.proto.ks.mode == KM_I64);  // a failed compile leaves the generic kernels in use
    return VR_OK;
}

vr_status vr_agg_memory(vr_agg* agg, int64_t* bytes) {
    if (!agg || !bytes) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    *bytes = (int64_t)(h.cells.bytes + h.state.bytes + h.counters.bytes + h.order.bytes + h.order_tmp.bytes +
                       h.cub_tmp.bytes + h.out_stage.bytes + h.row_slot.bytes + h.sort_keys.bytes +
                       h.sort_vals.bytes + h.sort_vals_in.bytes + h.firsts.bytes + h.firsts_sorted.bytes +
                       h.batch_descs.bytes + h.rgroups.bytes + h.resume.bytes + h.p_cand.bytes +
                       h.p_cand_count.bytes + h.p_hist.bytes + h.p_hot_keys.bytes + h.p_hot_part.bytes +
                       h.p_hot_total.bytes + h.p_pool.bytes + h.p_count.bytes + h.p_base.bytes + h.p_rtot.bytes + h.p_rstart.bytes +
                       h.p_resume.bytes + h.p_tile_batch.bytes + h.p_pending.bytes);
    for (auto& sk : h.strs) *bytes += sk.table.memory() + (int64_t)sk.remap.bytes;
    return VR_OK;
}

vr_status vr_agg_reset(vr_agg* agg) {
    if (!agg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    hipStream_t s = h.dev->stream;
    VR_TRY(enter(h, s));
    if (h.order_count >= 0 && (uint64_t)h.order_count < h.slots / 4) {
        // only the rows the last emission listed hold anything: reset those
        UpdateArgs a = h.proto;
        a.t = view_of(h);
        uint64_t empty = h.proto.ks.mode == KM_WIDE ? 0ull : EMPTY_KEY;
        if (h.order_count > 0)
            hipLaunchKernelGGL(reset_rows_kernel, dim3(grid_for(h.order_count, 256, h.dev->num_cus * 8)), dim3(256), 0, s,
                               h.order.as<uint64_t>(), h.order_count, h.cells.as<uint64_t>(), h.proto.ks.width, h.slots,
                               empty, a);
        VR_TRY_HIP(hipGetLastError());
        VR_TRY(clear_rgroups(h, h.rgroups, h.pbits, s));
    } else {
        VR_TRY(alloc_table(h, h.slots, h.pbits, h.cells, h.state, h.rgroups, s));
    }
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h.counters.as<uint64_t>(), (uint64_t)C_WORDS, 0ull);
    h.groups_upper = 0;
    h.rows_since_reset = 0;
    if (h.ungrouped) VR_TRY(insert_single_group(h, s));
    // no wait here: the next call on any stream is ordered after these fills (enter), and the
    // host can describe that call's batches while they run
    h.order_count = -1;
    return leave(h, s);
}

vr_status vr_agg_graph_create(vr_agg* agg, const vr_batch_set* bs, const vr_launch_config* cfg, int64_t max_groups,
                              vr_agg_graph** out) {
    if (!agg || !bs || !out) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    const int nb = bs->num_batches;
    if (nb <= 0) return fail(VR_ERR_INVALID_ARGUMENT, "a graph needs at least one batch");
    if (max_groups <= 0) return fail(VR_ERR_INVALID_ARGUMENT, "max_groups must be positive");
    if (h.ordered) return fail(VR_ERR_UNSUPPORTED, "updates that fold in row order after a sort are not captured");
    if (h.ungrouped) return fail(VR_ERR_UNSUPPORTED, "ungrouped aggregates are not captured");
    const int nk = (int)h.key_types.size(), nv = (int)h.value_types.size();
    if (bs->filter_col >= nv || (bs->filter_col >= 0 && h.value_types[bs->filter_col] != VR_INT32))
        return fail(VR_ERR_INVALID_ARGUMENT, "the filter column must be an INT32 value column");
    uint64_t all_rows = 0;
    for (int b = 0; b < nb; b++) {
        if (bs->lengths[b] < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative batch length");
        all_rows += (uint64_t)bs->lengths[b];
    }
    if (h.half_sums && all_rows >= (1ull << 32))
        return fail(VR_ERR_UNSUPPORTED, "an aggregate with integer or decimal sums takes fewer than 2^32 rows per step");
    for (int b = 0; b < nb; b++) {
        for (int c = 0; c < nk; c++) {
            const vr_column& k = bs->keys[(size_t)b * nk + c];
            if (k.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "a captured update reads device-resident batches");
            if (k.type == VR_UTF8 && (!k.dictionary || !bs->keys[c].dictionary || k.dictionary->data != bs->keys[c].dictionary->data))
                return fail(VR_ERR_UNSUPPORTED, "a captured update takes one dictionary per UTF8 key");
        }
        for (int c = 0; c < nv; c++)
            if (bs->values[(size_t)b * nv + c].mem != VR_MEM_DEVICE)
                return fail(VR_ERR_UNSUPPORTED, "a captured update reads device-resident batches");
    }
    std::unique_ptr<vr_agg_graph> g(new vr_agg_graph());
    g->owner.hold(agg->owner.dev);
    g->agg = agg;
    hipStream_t s = h.dev->stream;
    VR_TRY(enter(h, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    // the table sized for max_groups up front: nothing grows inside the graph (a step with more
    // groups than the table holds sets the table-full error, reported by the next read as OOM)
    const uint64_t want = next_pow2((uint64_t)((double)max_groups / h.max_load) + 1);
    if (want > h.slots) VR_TRY(grow_table(h, want, s));
    UpdateArgs a = h.proto;
    a.filter_col = bs->filter_col;
    a.filter_threshold = bs->filter_threshold;
    int grid, block;
    size_t lds_bytes;
    a.total_tiles = 1;
    launch_shape(h, cfg, 0, a, &grid, &block, &lds_bytes, true);
    const int64_t T = (int64_t)block * a.rows_per_thread;
    std::vector<BatchDesc> descs((size_t)nb);
    KeyIds ids(s);
    int64_t tiles = 0, rows = 0;
    for (int b = 0; b < nb; b++) {
        StagePlan plan;
        int32_t mem;
        const int64_t n = bs->lengths[b];
        VR_TRY(describe(h, bs->keys + (size_t)b * nk, bs->values + (size_t)b * nv, n, nullptr, descs[b], plan, ids.bufs,
                        &mem, s));
        descs[b].row_base = rows;  // every replay starts from a cleared table
        descs[b].tile_begin = tiles;
        tiles += (n + T - 1) / T;
        rows += n;
    }
    // the dictionary maps the descriptors point at, copied to the graph (the handle's own map of a
    // key changes when another dictionary is mapped)
    size_t remap_bytes = 0;
    for (int c = 0; c < nk; c++)
        if (h.key_types[c] == VR_UTF8) remap_bytes += align_up((size_t)descs[0].keys[c].remap_len * 4, 256);
    if (remap_bytes) {
        VR_TRY(g->remaps.ensure(h.dev, remap_bytes));
        size_t off = 0;
        for (int c = 0; c < nk; c++) {
            if (h.key_types[c] != VR_UTF8) continue;
            int32_t* dst = (int32_t*)(g->remaps.as<uint8_t>() + off);
            VR_TRY_HIP(hipMemcpyAsync(dst, descs[0].keys[c].remap, (size_t)descs[0].keys[c].remap_len * 4,
                                      hipMemcpyDeviceToDevice, s));
            for (int b = 0; b < nb; b++) descs[b].keys[c].remap = dst;
            off += align_up((size_t)descs[0].keys[c].remap_len * 4, 256);
        }
    }
    VR_TRY(g->descs.ensure(h.dev, sizeof(BatchDesc) * (size_t)nb));
    VR_TRY_HIP(hipMemcpyAsync(g->descs.p, descs.data(), sizeof(BatchDesc) * (size_t)nb, hipMemcpyHostToDevice, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    const int cap = batches_cap(h, cfg);
    grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, cap));
    a.batches = g->descs.as<BatchDesc>();
    a.num_batches = nb;
    a.tile_lo = 0;
    a.total_tiles = tiles;
    a.epoch = ++h.epoch;
    a.t = view_of(h);
    // capture: every node a kernel (no memset nodes), no read-back, no allocation
    hipStream_t cs = nullptr;
    VR_TRY_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const int W = h.proto.ks.width;
        const int fcap = h.dev->num_cus * 16;
        const uint64_t empty = h.proto.ks.mode == KM_WIDE ? 0ull : EMPTY_KEY;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(h.slots * W, 256, fcap)), dim3(256), 0, cs, h.cells.as<uint64_t>(),
                           (uint64_t)(h.slots * W), empty);
        UpdateArgs ia = h.proto;
        ia.t.stride = h.stride;
        hipLaunchKernelGGL(init_state_kernel, dim3(grid_for((h.slots + 3 + 63) / 64 * 64, 256, fcap)), dim3(256), 0, cs,
                           h.state.as<uint64_t>(), (uint64_t)(h.slots + 3), ia);
        const uint64_t rg_words = ((1ull << h.pbits) + 1) / 2;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(rg_words, 256, 64)), dim3(256), 0, cs, h.rgroups.as<uint64_t>(),
                           rg_words, 0ull);
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, cs, h.counters.as<uint64_t>(), (uint64_t)C_WORDS, 0ull);
        if (rows > 0) hipLaunchKernelGGL(agg_update_kernel<true>, dim3(grid), dim3(block), lds_bytes, cs, a);
        if (rows > 0 && h.has_f64_minmax)
            hipLaunchKernelGGL(nan_payload_kernel<true>, dim3(grid), dim3(block), 0, cs, a);
        hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return hip_fail(e, "capturing the batched update");
    VR_TRY_HIP(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
    g->rows = rows;
    g->cells = h.cells.p;
    g->state = h.state.p;
    g->counters = h.counters.p;
    g->slots = h.slots;
    VR_TRY(leave(h, s));
    agg->graphs.push_back(g.get());
    *out = g.release();
    return VR_OK;
}

vr_status vr_agg_graph_launch(vr_agg_graph* g, void* stream) {
    if (!g || !g->exec) return fail(VR_ERR_INVALID_ARGUMENT, "null graph");
    if (!g->agg) return fail(VR_ERR_STATE, "the aggregate the graph was captured from was destroyed");
    AggImpl& h = g->agg->h;
    if (h.cells.p != g->cells || h.state.p != g->state || h.counters.p != g->counters || h.slots != g->slots)
        return fail(VR_ERR_STATE, "the aggregate's table moved since the graph was captured (capture it again)");
    hipStream_t s = h.dev->pick(stream);
    VR_TRY(enter(h, s));
    VR_TRY_HIP(hipGraphLaunch(g->exec, s));
    VR_TRY_HIP(hipEventRecord(g->done, s));
    // the host's view of the table after the step: cleared, then folded over the graph's rows
    h.groups_upper = g->rows;
    h.rows_seen = g->rows;
    h.rows_since_reset = (uint64_t)g->rows;
    h.order_count = -1;
    h.dev->count(&vr_stats::calls, 1);
    h.dev->count(&vr_stats::input_rows, g->rows);
    h.dev->count(&vr_stats::kernel_launches, 1);  // one graph launch
    return leave(h, s);
}

vr_status vr_agg_graph_destroy(vr_agg_graph* g) {
    if (!g) return VR_OK;
    if (g->agg) {
        std::vector<vr_agg_graph*>& live = g->agg->graphs;
        live.erase(std::remove(live.begin(), live.end(), g), live.end());
    }
    delete g;
    return VR_OK;
}

vr_status vr_agg_set_timing(vr_agg* agg, int32_t enable) {
    if (!agg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    agg->h.timing = enable != 0;
    agg->h.timer_used = 0;
    return VR_OK;
}

vr_status vr_agg_kernel_time(vr_agg* agg, double* ms, int64_t* launches) {
    if (!agg) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    AggImpl& h = agg->h;
    double total = 0;
    for (size_t i = 0; i < h.timer_used; i++) {
        VR_TRY_HIP(hipEventSynchronize(h.timer_events[i].second));
        float t = 0;
        VR_TRY_HIP(hipEventElapsedTime(&t, h.timer_events[i].first, h.timer_events[i].second));
        total += t;
    }
    if (ms) *ms = total;
    if (launches) *launches = (int64_t)h.timer_used;
    h.timer_used = 0;
    return VR_OK;
}

vr_status vr_agg_destroy(vr_agg* agg) {
    if (!agg) return VR_OK;
    (void)hipStreamSynchronize(agg->h.dev->stream);
    if (agg->h.last_ev) (void)hipEventSynchronize(agg->h.last_ev);  // work queued on a caller's stream
    for (vr_agg_graph* g : agg->graphs) g->agg = nullptr;  // a graph outliving its aggregate stays destroyable
    delete agg;
    return VR_OK;
}

}  // extern "C"
