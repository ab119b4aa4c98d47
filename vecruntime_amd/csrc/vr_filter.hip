// Selection producers, compaction and gathers on gfx950: vr_filter_compare, vr_compact_utf8_bytes,
// vr_compact, vr_gather_utf8_bytes and vr_gather of include/vr.h.
//
// Replaces the filter operator's predicate-to-selection step and its compacting branch
// (spark/src/main/scala/org/apache/spark/sql/vecruntime/VectorFilterExec.scala:80-115) for a
// column-versus-literal comparison: CompareKernels.compareScalar
// (kernels/src/main/java/io/vecruntime/kernels/CompareKernels.java:57-160) + VectorExpr.selection
// (spark/src/main/scala/io/vecruntime/spark/expr/VectorExpr.scala:438-443), then
// CompactKernels.compactFixed / compactUtf8 (CompactKernels.java:108-471).
//
// Every kernel is a streaming pass bound by HBM: one wave per 64-row word (the selection, validity
// and BOOL bitmaps are 64-bit words, built with a ballot), coalesced lane reads, and compacted
// output positions from an exclusive scan of the words' popcounts.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "vr_device.h"
#include "vr_internal.h"

namespace vr {
namespace {

constexpr int FILTER_BLOCK = 256;  // four waves

__device__ __forceinline__ bool bit_of(const uint8_t* bm, int64_t i) { return (bm[i >> 3] >> (i & 7)) & 1u; }

__device__ __forceinline__ uint64_t word_at(const uint8_t* bm, int64_t w, int64_t n) {
    // the bits of rows [64w, min(64w + 64, n)) of an LSB-first bitmap; a null bitmap is all ones
    const int64_t base = w * 64;
    const int64_t left = n - base;
    const uint64_t live = left >= 64 ? ~0ull : ((1ull << left) - 1);
    if (!bm) return live;
    // a whole word of an 8-byte-aligned bitmap in one load; the last, partial word byte by byte
    // (input bitmaps hold ceil(n/8) bytes, nothing past them)
    if (left >= 64 && ((uintptr_t)bm & 7) == 0) return ((const uint64_t*)bm)[w];
    uint64_t x = 0;
    const int64_t bytes = (left + 7) / 8 < 8 ? (left + 7) / 8 : 8;
    for (int64_t b = 0; b < bytes; b++) x |= (uint64_t)bm[w * 8 + b] << (8 * b);
    return x & live;
}

// Three-way comparison of row i against the literal (CompareOp.nanSafeCompare for doubles,
// Decimal128.compare for the two-limb lane: signed high limbs, then unsigned low limbs).
__device__ __forceinline__ int compare_row(int32_t type, const void* data, int64_t i, const vr_scalar& s) {
    switch (type) {
        case VR_INT32: {
            const int32_t x = ((const int32_t*)data)[i], y = (int32_t)s.i64;
            return (x > y) - (x < y);
        }
        case VR_INT64: {
            const int64_t x = ((const int64_t*)data)[i], y = s.i64;
            return (x > y) - (x < y);
        }
        case VR_FLOAT64: {
            const double x = ((const double*)data)[i], y = s.f64;
            const bool xn = x != x, yn = y != y;
            if ((xn && yn) || x == y) return 0;
            if (xn) return 1;
            if (yn) return -1;
            return x > y ? 1 : -1;
        }
        default: {  // VR_DECIMAL128
            const int64_t* p = (const int64_t*)data + 2 * i;
            const int64_t xh = p[1], yh = s.hi;
            if (xh != yh) return xh > yh ? 1 : -1;
            const uint64_t xl = (uint64_t)p[0], yl = (uint64_t)s.lo;
            return (xl > yl) - (xl < yl);
        }
    }
}

__device__ __forceinline__ bool holds(int op, int c) {
    switch (op) {
        case VR_CMP_EQ: return c == 0;
        case VR_CMP_NE: return c != 0;
        case VR_CMP_LT: return c < 0;
        case VR_CMP_LE: return c <= 0;
        case VR_CMP_GT: return c > 0;
        default: return c >= 0;
    }
}

struct CompareArgs {
    const void* data;
    const uint8_t* valid;
    const uint8_t* sel;
    int64_t n;
    int32_t type, op;
    vr_scalar s;
    uint64_t* out;
    unsigned long long* partial;  // selected rows per workgroup (summed by sum_partials_kernel)
};

constexpr int FU = 4;    // 64-row words a compaction or gather wave has in flight per step
constexpr int FUC = 16;  // the same for a comparison (its wave holds only the values)

// The lane's value as a comparable pair (hi, lo): integers and doubles use one word.
template <int32_t TYPE>
struct Lane;
template <>
struct Lane<VR_INT32> {
    typedef int32_t T;
    __device__ static T load(const void* d, int64_t i) { return ((const int32_t*)d)[i]; }
    __device__ static int cmp(T x, const vr_scalar& s) {
        const int32_t y = (int32_t)s.i64;
        return (x > y) - (x < y);
    }
};
template <>
struct Lane<VR_INT64> {
    typedef int64_t T;
    __device__ static T load(const void* d, int64_t i) { return ((const int64_t*)d)[i]; }
    __device__ static int cmp(T x, const vr_scalar& s) { return (x > s.i64) - (x < s.i64); }
};
template <>
struct Lane<VR_FLOAT64> {
    typedef double T;
    __device__ static T load(const void* d, int64_t i) { return ((const double*)d)[i]; }
    __device__ static int cmp(T x, const vr_scalar& s) {
        const double y = s.f64;
        const bool xn = x != x, yn = y != y;
        if ((xn && yn) || x == y) return 0;
        if (xn) return 1;
        if (yn) return -1;
        return x > y ? 1 : -1;
    }
};
template <>
struct Lane<VR_DECIMAL128> {
    struct T {
        int64_t lo, hi;
    };
    __device__ static T load(const void* d, int64_t i) {
        const int64_t* p = (const int64_t*)d + 2 * i;
        return T{p[0], p[1]};
    }
    __device__ static int cmp(T x, const vr_scalar& s) {
        if (x.hi != s.hi) return x.hi > s.hi ? 1 : -1;
        return ((uint64_t)x.lo > (uint64_t)s.lo) - ((uint64_t)x.lo < (uint64_t)s.lo);
    }
};

// A wave takes FUC consecutive words per step: the values of all of them are loaded before any
// comparison, so each wave has FUC coalesced reads in flight.
template <int32_t TYPE>
__global__ __launch_bounds__(FILTER_BLOCK) void compare_kernel(CompareArgs a) {
    typedef Lane<TYPE> L;
    const int lane = threadIdx.x & 63;
    const int64_t words = (a.n + 63) / 64;
    const int64_t waves = (int64_t)gridDim.x * (FILTER_BLOCK / 64);
    unsigned long long total = 0;
    for (int64_t w0 = ((int64_t)blockIdx.x * (FILTER_BLOCK / 64) + threadIdx.x / 64) * FUC; w0 < words; w0 += waves * FUC) {
        typename L::T x[FUC];
#pragma unroll
        for (int u = 0; u < FUC; u++) {
            const int64_t i = (w0 + u) * 64 + lane;
            if (i < a.n) x[u] = L::load(a.data, i);
        }
#pragma unroll
        for (int u = 0; u < FUC; u++) {
            const int64_t w = w0 + u;
            if (w >= words) break;  // uniform over the wave
            const uint64_t live = word_at(a.sel, w, a.n) & word_at(a.valid, w, a.n);
            const bool on = ((live >> lane) & 1ull) && holds(a.op, L::cmp(x[u], a.s));
            const uint64_t bits = __ballot(on);
            if (lane == 0) {
                a.out[w] = bits;
                total += (unsigned long long)__popcll(bits);
            }
        }
    }
    // the workgroup's count, summed by one later workgroup (no contended atomics)
    __shared__ unsigned long long wave_total[FILTER_BLOCK / 64];
    if (lane == 0) wave_total[threadIdx.x / 64] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int q = 0; q < FILTER_BLOCK / 64; q++) t += wave_total[q];
        a.partial[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const unsigned long long* partial, int n, unsigned long long* out) {
    typedef hipcub::BlockReduce<unsigned long long, 256> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    unsigned long long t = 0;
    for (int q = threadIdx.x; q < n; q += 256) t += partial[q];
    const unsigned long long sum = Reduce(tmp).Sum(t);
    if (threadIdx.x == 0) *out = sum;
}

// Selected rows per 64-row word (the exclusive scan of these places each word's rows).
__global__ void word_counts_kernel(const uint8_t* sel, int64_t n, uint32_t* counts) {
    const int64_t words = (n + 63) / 64;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (int64_t)gridDim.x * blockDim.x)
        counts[w] = (uint32_t)__popcll(word_at(sel, w, n));
}

struct CompactCol {
    int32_t kind;        // 0: 4-byte lane, 1: 8-byte, 2: 16-byte, 3: BOOL bits, 4: plain UTF8
    const void* data;
    const uint8_t* valid;  // nullptr: no validity to carry
    const int32_t* offsets;
    void* out_data;
    uint64_t* out_valid;   // nullptr: not written
    int32_t* out_offsets;
    int32_t* lens;         // plain UTF8: the selected values' lengths, in compacted order
};

// ORs `bits` (the low `k` bits significant) into a bitmap at bit position `at`.
__device__ __forceinline__ void or_bits(uint64_t* bm, uint64_t at, uint64_t bits, int k) {
    if (!k) return;
    const uint64_t q = at >> 6;
    const int sh = (int)(at & 63);
    atomicOr((unsigned long long*)&bm[q], (unsigned long long)(bits << sh));
    if (sh && k > 64 - sh) atomicOr((unsigned long long*)&bm[q + 1], (unsigned long long)(bits >> (64 - sh)));
}

// The bits of the selected lanes, packed in lane order (a compress of `v` by `sel` over the wave).
__device__ __forceinline__ uint64_t packed_bits(bool selected, bool v, uint64_t sel_word, int lane) {
    const uint64_t below = sel_word & ((1ull << lane) - 1);
    uint64_t mine = (selected && v) ? (1ull << __popcll(below)) : 0ull;
    for (int d = 1; d < 64; d <<= 1) mine |= __shfl_xor(mine, d, 64);
    return mine;
}

// A wave takes FU consecutive words per step: their selection words, positions and the selected
// lanes' values are all loaded before anything is stored.
__global__ __launch_bounds__(FILTER_BLOCK) void compact_kernel(CompactCol c, const uint8_t* sel, int64_t n,
                                                            const uint32_t* word_off) {
    const int lane = threadIdx.x & 63;
    const int64_t words = (n + 63) / 64;
    const int64_t waves = (int64_t)gridDim.x * (FILTER_BLOCK / 64);
    for (int64_t w0 = ((int64_t)blockIdx.x * (FILTER_BLOCK / 64) + threadIdx.x / 64) * FU; w0 < words; w0 += waves * FU) {
        uint64_t sw[FU], base[FU], x0[FU], x1[FU];
        bool vb[FU];
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int64_t w = w0 + u;
            sw[u] = w < words ? word_at(sel, w, n) : 0ull;
            base[u] = w < words ? word_off[w] : 0;
            const int64_t i = w * 64 + lane;
            const bool selected = (sw[u] >> lane) & 1ull;
            x0[u] = x1[u] = 0;
            vb[u] = false;
            if (selected) {
                switch (c.kind) {
                    case 0: x0[u] = (uint32_t)((const int32_t*)c.data)[i]; break;
                    case 1: x0[u] = ((const uint64_t*)c.data)[i]; break;
                    case 2:
                        x0[u] = ((const uint64_t*)c.data)[2 * i];
                        x1[u] = ((const uint64_t*)c.data)[2 * i + 1];
                        break;
                    case 3: vb[u] = bit_of((const uint8_t*)c.data, i); break;
                    case 4: x0[u] = (uint32_t)(c.offsets[i + 1] - c.offsets[i]); break;
                    default: break;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < FU; u++) {
            if (!sw[u]) continue;  // uniform over the wave
            const int64_t i = (w0 + u) * 64 + lane;
            const bool selected = (sw[u] >> lane) & 1ull;
            const uint64_t o = base[u] + (uint64_t)__popcll(sw[u] & ((1ull << lane) - 1));
            const int k = __popcll(sw[u]);
            if (selected) {
                switch (c.kind) {
                    case 0: ((uint32_t*)c.out_data)[o] = (uint32_t)x0[u]; break;
                    case 1: ((uint64_t*)c.out_data)[o] = x0[u]; break;
                    case 2:
                        ((uint64_t*)c.out_data)[2 * o] = x0[u];
                        ((uint64_t*)c.out_data)[2 * o + 1] = x1[u];
                        break;
                    case 4: c.lens[o] = (int32_t)x0[u]; break;  // lengths, scanned into offsets
                    default: break;
                }
            }
            if (c.kind == 3) {
                const uint64_t bits = packed_bits(selected, vb[u], sw[u], lane);
                if (lane == 0) or_bits((uint64_t*)c.out_data, base[u], bits, k);
            }
            if (c.out_valid) {
                const bool v = selected && (!c.valid || bit_of(c.valid, i));
                const uint64_t bits = packed_bits(selected, v, sw[u], lane);
                if (lane == 0) or_bits(c.out_valid, base[u], bits, k);
            }
        }
    }
}

// Copies the bytes of the selected plain UTF8 values to their compacted offsets.
__global__ void utf8_bytes_kernel(CompactCol c, const uint8_t* sel, int64_t n, const uint32_t* word_off) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = i >> 6;
        const uint64_t sw = word_at(sel, w, n);
        const int lane = (int)(i & 63);
        if (!((sw >> lane) & 1ull)) continue;
        const uint64_t o = (uint64_t)word_off[w] + (uint64_t)__popcll(sw & ((1ull << lane) - 1));
        const int32_t src = c.offsets[i], len = c.offsets[i + 1] - src, dst = c.out_offsets[o];
        const uint8_t* s = (const uint8_t*)c.data + src;
        uint8_t* d = (uint8_t*)c.out_data + dst;
        for (int32_t b = 0; b < len; b++) d[b] = s[b];
    }
}

__global__ void selected_bytes_kernel(const int32_t* offsets, const uint8_t* sel, int64_t n, unsigned long long* total) {
    unsigned long long t = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!sel || ((sel[i >> 3] >> (i & 7)) & 1u)) t += (unsigned long long)(offsets[i + 1] - offsets[i]);
    if (t) atomicAdd(total, t);
}

struct GatherArgs {
    int32_t kind;  // 0: 4-byte lane, 1: 8-byte, 2: 16-byte, 3: BOOL bits, 4: plain UTF8
    const void* data;
    const uint8_t* valid;
    const int32_t* offsets;
    const int32_t* idx;
    int64_t m;
    void* out_data;
    uint64_t* out_valid;
    int32_t* out_offsets;
    int32_t* lens;
};

// GatherKernels.gatherFixed / gatherBits / gatherValidity (GatherKernels.java:99-248): a wave per
// 64 output rows, so bitmaps are written as whole words.
__global__ __launch_bounds__(FILTER_BLOCK) void gather_kernel(GatherArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t words = (a.m + 63) / 64;
    const int64_t waves = (int64_t)gridDim.x * (FILTER_BLOCK / 64);
    for (int64_t w0 = ((int64_t)blockIdx.x * (FILTER_BLOCK / 64) + threadIdx.x / 64) * FU; w0 < words; w0 += waves * FU) {
        int64_t idx[FU];
        uint64_t x0[FU], x1[FU];
        bool valid[FU], bitv[FU];
#pragma unroll
        for (int u = 0; u < FU; u++) {  // the indices of FU words, then their rows' values
            const int64_t o = (w0 + u) * 64 + lane;
            idx[u] = o < a.m ? a.idx[o] : -1;
        }
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int64_t i = idx[u];
            const bool real = i >= 0;
            valid[u] = real && (!a.valid || bit_of(a.valid, i));
            x0[u] = x1[u] = 0;
            bitv[u] = false;
            if (real) {
                switch (a.kind) {
                    case 0: x0[u] = (uint32_t)((const int32_t*)a.data)[i]; break;
                    case 1: x0[u] = ((const uint64_t*)a.data)[i]; break;
                    case 2:
                        x0[u] = ((const uint64_t*)a.data)[2 * i];
                        x1[u] = ((const uint64_t*)a.data)[2 * i + 1];
                        break;
                    case 3: bitv[u] = bit_of((const uint8_t*)a.data, i); break;
                    case 4: x0[u] = valid[u] ? (uint32_t)(a.offsets[i + 1] - a.offsets[i]) : 0u; break;
                    default: break;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int64_t w = w0 + u;
            if (w >= words) break;  // uniform over the wave
            const int64_t o = w * 64 + lane;
            if (o < a.m) {
                switch (a.kind) {
                    case 0: ((uint32_t*)a.out_data)[o] = (uint32_t)x0[u]; break;
                    case 1: ((uint64_t*)a.out_data)[o] = x0[u]; break;
                    case 2:
                        ((uint64_t*)a.out_data)[2 * o] = x0[u];
                        ((uint64_t*)a.out_data)[2 * o + 1] = x1[u];
                        break;
                    case 4: a.lens[o] = (int32_t)x0[u]; break;
                    default: break;
                }
            }
            if (a.kind == 3) {
                const uint64_t bits = __ballot(bitv[u]);
                if (lane == 0) ((uint64_t*)a.out_data)[w] = bits;
            }
            if (a.out_valid) {
                const uint64_t bits = __ballot(valid[u]);
                if (lane == 0) a.out_valid[w] = bits;
            }
        }
    }
}

// GatherKernels.gatherUtf8 (:274-308): the bytes of the valid, real rows at their offsets.
__global__ void gather_utf8_kernel(GatherArgs a) {
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < a.m; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = a.idx[o];
        if (i < 0 || (a.valid && !bit_of(a.valid, i))) continue;
        const int32_t src = a.offsets[i], len = a.offsets[i + 1] - src, dst = a.out_offsets[o];
        const uint8_t* s = (const uint8_t*)a.data + src;
        uint8_t* d = (uint8_t*)a.out_data + dst;
        for (int32_t b = 0; b < len; b++) d[b] = s[b];
    }
}

// GatherKernels.gatherUtf8Bytes (:253-268).
__global__ void gather_bytes_kernel(const int32_t* offsets, const uint8_t* valid, const int32_t* idx, int64_t m,
                                    unsigned long long* total) {
    unsigned long long t = 0;
    for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < m; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx[o];
        if (i >= 0 && (!valid || bit_of(valid, i))) t += (unsigned long long)(offsets[i + 1] - offsets[i]);
    }
    if (t) atomicAdd(total, t);
}

__global__ void store_i32_kernel(int32_t* p, int32_t v) {
    if (threadIdx.x == 0) *p = v;
}

__global__ void zero_words_kernel(uint64_t* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0;
}

int grid_of(uint64_t work, int block, int cap) {
    const uint64_t g = (work + block - 1) / block;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)cap));
}

vr_status check_device(const vr_column& c) {
    if (c.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "filters take device-resident columns");
    return VR_OK;
}

}  // namespace
}  // namespace vr

using namespace vr;

extern "C" {

vr_status vr_filter_compare(vr_device* dev, const vr_column* col, int32_t op, const vr_scalar* scalar,
                            const uint8_t* selection, int64_t n, uint8_t* out, int64_t* count, void* stream) {
    if (!dev || !col || !scalar || (!out && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (op < VR_CMP_EQ || op > VR_CMP_GE) return fail(VR_ERR_INVALID_ARGUMENT, "comparison operator");
    if (n < 0 || n > col->length) return fail(VR_ERR_INVALID_ARGUMENT, "row count outside the column");
    if (col->type != VR_INT32 && col->type != VR_INT64 && col->type != VR_FLOAT64 && col->type != VR_DECIMAL128)
        return fail(VR_ERR_UNSUPPORTED, "comparisons over this lane run on the CPU kernels");
    VR_TRY(check_device(*col));
    if (count) *count = 0;
    if (n == 0) return VR_OK;
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    const uint64_t words = (uint64_t)(n + 63) / 64;
    const dim3 grid(grid_of((words + FUC - 1) / FUC * 64, FILTER_BLOCK, d.num_cus * 32));
    DevBuf cnt;
    VR_TRY(cnt.ensure(&d, ((size_t)grid.x + 1) * 8));
    CompareArgs a{col->data, col->validity, selection, n, col->type, op, *scalar, (uint64_t*)out,
                  cnt.as<unsigned long long>() + 1};
    switch (col->type) {
        case VR_INT32: hipLaunchKernelGGL(compare_kernel<VR_INT32>, grid, dim3(FILTER_BLOCK), 0, s, a); break;
        case VR_INT64: hipLaunchKernelGGL(compare_kernel<VR_INT64>, grid, dim3(FILTER_BLOCK), 0, s, a); break;
        case VR_FLOAT64: hipLaunchKernelGGL(compare_kernel<VR_FLOAT64>, grid, dim3(FILTER_BLOCK), 0, s, a); break;
        default: hipLaunchKernelGGL(compare_kernel<VR_DECIMAL128>, grid, dim3(FILTER_BLOCK), 0, s, a); break;
    }
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, cnt.as<unsigned long long>() + 1, (int)grid.x,
                       cnt.as<unsigned long long>());
    VR_TRY_HIP(hipGetLastError());
    uint64_t c = 0;
    VR_TRY_HIP(hipMemcpyAsync(&c, cnt.p, 8, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (count) *count = (int64_t)c;
    d.count(&vr_stats::calls, 1);
    d.count(&vr_stats::input_rows, n);
    d.count(&vr_stats::kernel_launches, 2);
    return VR_OK;
}

vr_status vr_compact_utf8_bytes(vr_device* dev, const vr_column* col, int64_t n, const uint8_t* selection,
                                int64_t* bytes, void* stream) {
    if (!dev || !col || !bytes) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (col->type != VR_UTF8 || col->dictionary || !col->offsets)
        return fail(VR_ERR_INVALID_ARGUMENT, "expected a plain UTF8 column");
    if (n < 0 || n > col->length) return fail(VR_ERR_INVALID_ARGUMENT, "row count outside the column");
    VR_TRY(check_device(*col));
    *bytes = 0;
    if (n == 0) return VR_OK;
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    DevBuf tot;
    VR_TRY(tot.ensure(&d, 8));
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, tot.as<uint64_t>(), 1ull);
    hipLaunchKernelGGL(selected_bytes_kernel, dim3(grid_of((uint64_t)n, 256, d.num_cus * 16)), dim3(256), 0, s, col->offsets,
                       selection, n, tot.as<unsigned long long>());
    VR_TRY_HIP(hipGetLastError());
    uint64_t b = 0;
    VR_TRY_HIP(hipMemcpyAsync(&b, tot.p, 8, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    *bytes = (int64_t)b;
    return VR_OK;
}

vr_status vr_compact(vr_device* dev, const vr_column* columns, int32_t num_columns, int64_t n, const uint8_t* selection,
                     vr_out_column* out, int64_t* rows, void* stream) {
    if (!dev || (num_columns && (!columns || !out)) || num_columns < 0) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 0 || n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "compaction of batches past 2^31 rows");
    if (rows) *rows = 0;
    std::vector<CompactCol> cols(num_columns);
    for (int c = 0; c < num_columns; c++) {  // every column checked before anything is queued
        const vr_column& in = columns[c];
        const vr_out_column& o = out[c];
        VR_TRY(check_device(in));
        if (o.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "compaction writes device-resident outputs");
        if (in.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "column shorter than the batch");
        const bool dict = in.type == VR_UTF8 && in.dictionary;
        CompactCol& cc = cols[c];
        cc.data = in.data;
        cc.valid = in.validity;
        cc.out_data = o.data;
        cc.out_valid = (uint64_t*)o.validity;
        switch (dict ? VR_INT32 : in.type) {
            case VR_INT32: cc.kind = 0; break;
            case VR_INT64:
            case VR_FLOAT64: cc.kind = 1; break;
            case VR_DECIMAL128: cc.kind = 2; break;
            case VR_BOOL: cc.kind = 3; break;
            default: cc.kind = 4; break;  // plain UTF8
        }
        if (o.type != (dict ? VR_INT32 : in.type))
            return fail(VR_ERR_INVALID_ARGUMENT, "output lane differs from the column (a dictionary column compacts its int32 codes)");
        if (cc.kind == 4) {
            if (!in.offsets || !o.offsets) return fail(VR_ERR_INVALID_ARGUMENT, "a UTF8 compaction needs offsets");
            cc.offsets = in.offsets;
            cc.out_offsets = o.offsets;
        }
    }
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    const int cap = d.num_cus * 16;
    // the compacted position of each word's first selected row; every temporary lives until the
    // final synchronization
    const uint64_t words = (uint64_t)(n + 63) / 64;
    DevBuf counts, offs, tmp;
    std::vector<DevBuf> keep;
    VR_TRY(counts.ensure(&d, (words + 1) * 4));
    VR_TRY(offs.ensure(&d, (words + 1) * 4));
    uint32_t last[2] = {0, 0};
    if (words) {
        hipLaunchKernelGGL(word_counts_kernel, dim3(grid_of(words, 256, cap)), dim3(256), 0, s, selection, n, counts.as<uint32_t>());
        size_t tb = 0;
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<uint32_t>(), offs.as<uint32_t>(), (int)words, s));
        VR_TRY(tmp.ensure(&d, tb));
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, counts.as<uint32_t>(), offs.as<uint32_t>(), (int)words, s));
        VR_TRY_HIP(hipMemcpyAsync(&last[0], offs.as<uint32_t>() + words - 1, 4, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipMemcpyAsync(&last[1], counts.as<uint32_t>() + words - 1, 4, hipMemcpyDeviceToHost, s));
    }
    VR_TRY_HIP(hipStreamSynchronize(s));
    const int64_t m = (int64_t)last[0] + last[1];
    if (rows) *rows = m;
    for (int c = 0; c < num_columns; c++)
        if (out[c].capacity < m) return fail(VR_ERR_CAPACITY, "compacted columns need " + std::to_string(m) + " rows");
    const uint64_t mw = (uint64_t)(m + 63) / 64;
    for (int c = 0; c < num_columns; c++) {
        CompactCol& cc = cols[c];
        // bits are OR-ed in: clear the words the compacted rows cover
        if (cc.out_valid && mw)
            hipLaunchKernelGGL(zero_words_kernel, dim3(grid_of(mw, 256, cap)), dim3(256), 0, s, cc.out_valid, mw);
        if (cc.kind == 3 && mw)
            hipLaunchKernelGGL(zero_words_kernel, dim3(grid_of(mw, 256, cap)), dim3(256), 0, s, (uint64_t*)cc.out_data, mw);
        if (cc.kind == 4) {
            hipLaunchKernelGGL(store_i32_kernel, dim3(1), dim3(64), 0, s, cc.out_offsets, 0);
            keep.emplace_back();
            VR_TRY(keep.back().ensure(&d, (size_t)std::max<int64_t>(m, 1) * 4));
            cc.lens = keep.back().as<int32_t>();
        }
        if (words)
            hipLaunchKernelGGL(compact_kernel, dim3(grid_of((words + FU - 1) / FU * 64, FILTER_BLOCK, 2 * cap)), dim3(FILTER_BLOCK), 0,
                               s, cc, selection, n, offs.as<uint32_t>());
        if (cc.kind == 4 && m) {
            // lengths -> offsets[1..m] (running ends), then the bytes
            size_t tb = 0;
            VR_TRY_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, cc.lens, cc.out_offsets + 1, (int)m, s));
            keep.emplace_back();
            VR_TRY(keep.back().ensure(&d, tb));
            VR_TRY_HIP(hipcub::DeviceScan::InclusiveSum(keep.back().p, tb, cc.lens, cc.out_offsets + 1, (int)m, s));
            hipLaunchKernelGGL(utf8_bytes_kernel, dim3(grid_of((uint64_t)n, 256, cap)), dim3(256), 0, s, cc, selection, n,
                               offs.as<uint32_t>());
        }
        VR_TRY_HIP(hipGetLastError());
    }
    VR_TRY_HIP(hipStreamSynchronize(s));
    d.count(&vr_stats::calls, 1);
    d.count(&vr_stats::input_rows, n);
    d.count(&vr_stats::output_rows, m);
    return VR_OK;
}

vr_status vr_gather_utf8_bytes(vr_device* dev, const vr_column* col, const int32_t* idx, int64_t m, int64_t* bytes,
                               void* stream) {
    if (!dev || !col || !bytes || (!idx && m)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (col->type != VR_UTF8 || col->dictionary || !col->offsets) return fail(VR_ERR_INVALID_ARGUMENT, "expected a plain UTF8 column");
    VR_TRY(check_device(*col));
    *bytes = 0;
    if (m <= 0) return VR_OK;
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    DevBuf tot;
    VR_TRY(tot.ensure(&d, 8));
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, tot.as<uint64_t>(), 1ull);
    hipLaunchKernelGGL(gather_bytes_kernel, dim3(grid_of((uint64_t)m, 256, d.num_cus * 16)), dim3(256), 0, s, col->offsets,
                       col->validity, idx, m, tot.as<unsigned long long>());
    VR_TRY_HIP(hipGetLastError());
    uint64_t b = 0;
    VR_TRY_HIP(hipMemcpyAsync(&b, tot.p, 8, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    *bytes = (int64_t)b;
    return VR_OK;
}

vr_status vr_gather(vr_device* dev, const vr_column* col, const int32_t* idx, int64_t m, vr_out_column* out, void* stream) {
    if (!dev || !col || !out || (!idx && m)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (m < 0 || m >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "gathers of past 2^31 rows");
    VR_TRY(check_device(*col));
    if (out->mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "gathers write device-resident outputs");
    if (out->capacity < m) return fail(VR_ERR_CAPACITY, "the gather needs " + std::to_string(m) + " rows");
    if (col->type == VR_UTF8 && col->dictionary) return fail(VR_ERR_UNSUPPORTED, "gather a dictionary column's int32 codes");
    if (out->type != col->type) return fail(VR_ERR_INVALID_ARGUMENT, "output lane differs from the column");
    GatherArgs a{};
    switch (col->type) {
        case VR_INT32: a.kind = 0; break;
        case VR_INT64:
        case VR_FLOAT64: a.kind = 1; break;
        case VR_DECIMAL128: a.kind = 2; break;
        case VR_BOOL: a.kind = 3; break;
        default:
            a.kind = 4;
            if (!col->offsets || !out->offsets) return fail(VR_ERR_INVALID_ARGUMENT, "a UTF8 gather needs offsets");
    }
    a.data = col->data;
    a.valid = col->validity;
    a.offsets = col->offsets;
    a.idx = idx;
    a.m = m;
    a.out_data = out->data;
    a.out_valid = (uint64_t*)out->validity;
    a.out_offsets = out->offsets;
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    const int cap = d.num_cus * 16;
    DevBuf lens, tmp;
    if (a.kind == 4) {
        VR_TRY(lens.ensure(&d, (size_t)std::max<int64_t>(m, 1) * 4));
        a.lens = lens.as<int32_t>();
        hipLaunchKernelGGL(store_i32_kernel, dim3(1), dim3(64), 0, s, out->offsets, 0);
    }
    const uint64_t words = (uint64_t)(m + 63) / 64;
    if (words)
        hipLaunchKernelGGL(gather_kernel, dim3(grid_of((words + FU - 1) / FU * 64, FILTER_BLOCK, 2 * cap)), dim3(FILTER_BLOCK), 0, s, a);
    if (a.kind == 4 && m) {
        size_t tb = 0;
        VR_TRY_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, a.lens, out->offsets + 1, (int)m, s));
        VR_TRY(tmp.ensure(&d, tb));
        VR_TRY_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, a.lens, out->offsets + 1, (int)m, s));
        hipLaunchKernelGGL(gather_utf8_kernel, dim3(grid_of((uint64_t)m, 256, cap)), dim3(256), 0, s, a);
    }
    VR_TRY_HIP(hipGetLastError());
    VR_TRY_HIP(hipStreamSynchronize(s));
    d.count(&vr_stats::calls, 1);
    d.count(&vr_stats::output_rows, m);
    return VR_OK;
}

}  // extern "C"
