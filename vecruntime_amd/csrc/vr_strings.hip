// Device dictionary of byte strings: StrTable (vr_strings.h).
#include <algorithm>
#include <cstring>

#include "vr_device.h"
#include "vr_strings.h"

namespace vr {

namespace {

constexpr uint32_t PENDING = 0x80000000u;  // ref of a slot claimed by an input row, id not given yet
constexpr int32_t ROW_UNKNOWN = -1, ROW_NULL = -2;

struct StrView {
    uint64_t* cells;
    uint64_t mask;
    uint64_t* spans;  // per id: (arena offset << 32) | length
    uint8_t* arena;
};

// 64-bit hash of a byte string: little-endian 8-byte words folded through fmix64, the length
// mixed into the seed and the last word.
__device__ __forceinline__ uint64_t str_hash(const uint8_t* p, int32_t len) {
    uint64_t h = 0x243F6A8885A308D3ull ^ ((uint64_t)(uint32_t)len * 0x9E3779B97F4A7C15ull);
    int32_t k = 0;
    for (; k + 8 <= len; k += 8) {
        uint64_t w = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) w |= (uint64_t)p[k + j] << (8 * j);
        h = fmix64(h ^ w) * 0xC2B2AE3D27D4EB4Full;
    }
    uint64_t w = (uint64_t)(uint32_t)len << 56;
    for (int j = 0; k + j < len; j++) w ^= (uint64_t)p[k + j] << (8 * j);
    return fmix64(h ^ w);
}

__device__ __forceinline__ uint32_t str_fp(uint64_t h) { return (uint32_t)(h >> 33) | 1u; }

__device__ __forceinline__ bool same_bytes(const uint8_t* a, const uint8_t* b, int32_t len) {
    for (int32_t k = 0; k < len; k++)
        if (a[k] != b[k]) return false;
    return true;
}

// Row i's byte range, kept inside [begin, end) whatever the offsets hold (Arrow offsets never
// decrease; a column whose offsets do reads as shorter strings, and its insert fails).
__device__ __forceinline__ void row_range(const StrRows& r, int64_t i, int32_t* s0, int32_t* len) {
    const int32_t a = min(max(r.offsets[i], r.begin), r.end);
    const int32_t b = min(max(r.offsets[i + 1], a), r.end);
    *s0 = a;
    *len = b - a;
}

// out[i] = the slot holding row i's string (claimed for it when inserting), ROW_UNKNOWN when it is
// absent and not inserted, ROW_NULL for a null row.
__global__ void str_probe_kernel(StrView t, StrRows r, int insert, int32_t* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r.n; i += (int64_t)gridDim.x * blockDim.x) {
        if (r.valid && !bit_at(r.valid, i)) {
            out[i] = ROW_NULL;
            continue;
        }
        int32_t s0, len;
        row_range(r, i, &s0, &len);
        const uint8_t* p = r.bytes + s0;
        const uint64_t h = str_hash(p, len);
        const uint32_t fp = str_fp(h);
        uint64_t pos = h & t.mask;
        int32_t res = ROW_UNKNOWN;
        while (true) {
            uint64_t c = ld_coherent(&t.cells[pos]);
            if (c == 0) {
                if (!insert) break;
                c = atomic_cas_u64(&t.cells[pos], 0ull, ((uint64_t)fp << 32) | PENDING | (uint64_t)i);
                if (c == 0) {
                    res = (int32_t)pos;
                    break;
                }
            }
            if ((uint32_t)(c >> 32) == fp) {
                const uint32_t ref = (uint32_t)c;
                bool eq;
                if (ref & PENDING) {
                    const int64_t q = ref & ~PENDING;
                    int32_t q0, qlen;
                    row_range(r, q, &q0, &qlen);
                    eq = qlen == len && same_bytes(r.bytes + q0, p, len);
                } else {
                    const uint64_t sp = t.spans[ref];
                    eq = (int32_t)(uint32_t)sp == len && same_bytes(t.arena + (sp >> 32), p, len);
                }
                if (eq) {
                    res = (int32_t)pos;
                    break;
                }
            }
            pos = (pos + 1) & t.mask;
        }
        out[i] = res;
    }
}

// The rows that claimed a slot give their string an id and a place in the arena.
__global__ void str_assign_kernel(StrView t, StrRows r, const int32_t* out, uint64_t* counters, uint64_t arena_cap) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r.n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t pos = out[i];
        if (pos < 0) continue;
        const uint64_t c = t.cells[pos];
        if ((uint32_t)c != (PENDING | (uint32_t)i)) continue;
        int32_t s0, len;
        row_range(r, i, &s0, &len);
        const uint64_t id = atomic_add_u64(&counters[0], 1);
        const uint64_t off = atomic_add_u64(&counters[1], (uint64_t)len);
        if (off + (uint64_t)len > arena_cap) {  // only offsets that decrease overlap this far: refused
            atomic_add_u64(&counters[2], 1);
            t.spans[id] = 0;
            t.cells[pos] = (c & 0xFFFFFFFF00000000ull) | id;
            continue;
        }
        for (int32_t k = 0; k < len; k++) t.arena[off + k] = r.bytes[s0 + k];
        t.spans[id] = (off << 32) | (uint32_t)len;
        t.cells[pos] = (c & 0xFFFFFFFF00000000ull) | id;
    }
}

__global__ void str_ids_kernel(const uint64_t* cells, const int32_t* out, int64_t n, int32_t null_id, int32_t* ids) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t pos = out[i];
        ids[i] = pos >= 0 ? (int32_t)(uint32_t)cells[pos] : pos == ROW_NULL ? null_id : -1;
    }
}

// Every id's string into a larger, empty slot array.
__global__ void str_rehash_kernel(StrView t, int64_t count) {
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < count; id += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t sp = t.spans[id];
        const uint64_t h = str_hash(t.arena + (sp >> 32), (int32_t)(uint32_t)sp);
        const uint64_t v = ((uint64_t)str_fp(h) << 32) | (uint64_t)id;
        for (uint64_t pos = h & t.mask;; pos = (pos + 1) & t.mask)
            if (atomic_cas_u64(&t.cells[pos], 0ull, v) == 0) break;
    }
}

__global__ void str_fill_kernel(uint64_t* p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void str_set3_kernel(uint64_t* p, uint64_t a, uint64_t b) {
    if (threadIdx.x == 0) {
        p[0] = a;
        p[1] = b;
        p[2] = 0;
    }
}

uint64_t pow2_at_least(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

int grid_for_rows(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

// Grows `b` to at least `want` bytes keeping its first `keep` bytes.
vr_status grow_keep(Device* dev, DevBuf& b, size_t want, size_t keep, hipStream_t s) {
    if (want <= b.bytes && b.p) return VR_OK;
    DevBuf nb;
    VR_TRY(nb.ensure(dev, std::max(want, b.bytes * 2)));
    if (keep) VR_TRY_HIP(hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, s));
    VR_TRY_HIP(hipStreamSynchronize(s));  // the old block is read by the copy
    b = std::move(nb);
    return VR_OK;
}

}  // namespace

vr_status StrTable::reserve(Device* dev, int64_t more_strings, uint64_t more_bytes, hipStream_t s) {
    const uint64_t want_slots = std::max<uint64_t>(1024, pow2_at_least(2 * (uint64_t)(count_ + more_strings) + 2));
    VR_TRY(grow_keep(dev, spans_, (size_t)(count_ + more_strings + 1) * 8, (size_t)count_ * 8, s));
    VR_TRY(grow_keep(dev, arena_, std::max<size_t>(arena_used_ + more_bytes + 8, 1 << 16), arena_used_, s));
    if (!counters_.p) VR_TRY(counters_.ensure(dev, 64));
    if (want_slots > slots_) {
        VR_TRY(cells_.ensure(dev, want_slots * 8));
        hipLaunchKernelGGL(str_fill_kernel, dim3(grid_for_rows((int64_t)want_slots)), dim3(256), 0, s, cells_.as<uint64_t>(),
                           want_slots, 0ull);
        slots_ = want_slots;
        StrView v{cells_.as<uint64_t>(), slots_ - 1, spans_.as<uint64_t>(), arena_.as<uint8_t>()};
        if (count_) hipLaunchKernelGGL(str_rehash_kernel, dim3(grid_for_rows(count_)), dim3(256), 0, s, v, count_);
        VR_TRY_HIP(hipGetLastError());
    }
    return VR_OK;
}

vr_status StrTable::map(Device* dev, const StrRows& in, bool insert, int32_t null_id, int32_t* ids, hipStream_t s) {
    if (in.n <= 0) return VR_OK;
    if (in.n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "a UTF8 column of 2^31 rows or more");
    StrRows rows = in;
    int32_t ends[2] = {0, 0};
    VR_TRY_HIP(hipMemcpyAsync(&ends[0], rows.offsets, 4, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipMemcpyAsync(&ends[1], rows.offsets + rows.n, 4, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (ends[1] < ends[0]) return fail(VR_ERR_INVALID_ARGUMENT, "UTF8 offsets decrease");
    rows.begin = ends[0];
    rows.end = ends[1];
    if (insert) {
        if (arena_used_ + (uint64_t)(ends[1] - ends[0]) >= (1ull << 32))
            return fail(VR_ERR_UNSUPPORTED, "UTF8 key dictionaries past 4 GiB of bytes");
        VR_TRY(reserve(dev, rows.n, (uint64_t)(ends[1] - ends[0]), s));
    } else if (!slots_) {  // nothing inserted yet: every row is unknown (or null)
        VR_TRY(reserve(dev, 0, 0, s));
    }
    VR_TRY(scratch_.ensure(dev, (size_t)rows.n * 4));
    int32_t* out = scratch_.as<int32_t>();
    StrView v{cells_.as<uint64_t>(), slots_ - 1, spans_.as<uint64_t>(), arena_.as<uint8_t>()};
    const int grid = grid_for_rows(rows.n);
    hipLaunchKernelGGL(str_probe_kernel, dim3(grid), dim3(256), 0, s, v, rows, insert ? 1 : 0, out);
    if (insert) {
        hipLaunchKernelGGL(str_set3_kernel, dim3(1), dim3(64), 0, s, counters_.as<uint64_t>(), (uint64_t)count_, arena_used_);
        hipLaunchKernelGGL(str_assign_kernel, dim3(grid), dim3(256), 0, s, v, rows, out, counters_.as<uint64_t>(),
                           (uint64_t)arena_.bytes);
    }
    hipLaunchKernelGGL(str_ids_kernel, dim3(grid), dim3(256), 0, s, cells_.as<uint64_t>(), out, rows.n, null_id, ids);
    VR_TRY_HIP(hipGetLastError());
    dev->count(&vr_stats::kernel_launches, insert ? 4 : 2);
    if (insert) {
        uint64_t c[3] = {0, 0, 0};
        VR_TRY_HIP(hipMemcpyAsync(c, counters_.p, 24, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipStreamSynchronize(s));
        count_ = (int64_t)c[0];
        arena_used_ = std::min<uint64_t>(c[1], arena_.bytes);
        if (c[2]) return fail(VR_ERR_INVALID_ARGUMENT, "UTF8 offsets decrease (the column's strings overlap)");
    } else {
        VR_TRY_HIP(hipStreamSynchronize(s));  // the scratch is the table's: the next call rewrites it
    }
    return VR_OK;
}

vr_status StrTable::map_host(Device* dev, const int32_t* offsets, const uint8_t* bytes, const uint8_t* valid, int64_t n,
                             bool insert, int32_t null_id, int32_t* ids, hipStream_t s) {
    if (n <= 0) return VR_OK;
    const int32_t b0 = offsets[0], b1 = offsets[n];
    if (b1 < b0) return fail(VR_ERR_INVALID_ARGUMENT, "UTF8 offsets decrease");
    DevBuf doff, dbytes, dvalid;
    VR_TRY(doff.ensure(dev, (size_t)(n + 1) * 4));
    VR_TRY(dbytes.ensure(dev, (size_t)(b1 - b0) + 8));
    // offsets rebased to the copied bytes
    std::vector<int32_t> rebased((size_t)n + 1);
    for (int64_t i = 0; i <= n; i++) rebased[(size_t)i] = offsets[i] - b0;
    VR_TRY_HIP(hipMemcpyAsync(doff.p, rebased.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice, s));
    if (b1 > b0) VR_TRY_HIP(hipMemcpyAsync(dbytes.p, bytes + b0, (size_t)(b1 - b0), hipMemcpyHostToDevice, s));
    if (valid) {
        VR_TRY(dvalid.ensure(dev, bitmap_bytes(n)));
        VR_TRY_HIP(hipMemcpyAsync(dvalid.p, valid, bitmap_bytes(n), hipMemcpyHostToDevice, s));
    }
    dev->count(&vr_stats::h2d_bytes, (int64_t)((size_t)(n + 1) * 4 + (size_t)(b1 - b0)));
    StrRows r{doff.as<int32_t>(), dbytes.as<uint8_t>(), valid ? dvalid.as<uint8_t>() : nullptr, n};
    VR_TRY(map(dev, r, insert, null_id, ids, s));  // returns after its kernels: the copies may go
    return VR_OK;
}

vr_status StrTable::read(std::vector<int32_t>* offsets, std::vector<uint8_t>* bytes, hipStream_t s) const {
    if (arena_used_ >= (1ull << 31)) return fail(VR_ERR_UNSUPPORTED, "a dictionary of 2 GiB of bytes or more has no int32 offsets");
    std::vector<uint64_t> spans((size_t)count_);
    std::vector<uint8_t> arena((size_t)arena_used_);
    if (count_) VR_TRY_HIP(hipMemcpyAsync(spans.data(), spans_.p, (size_t)count_ * 8, hipMemcpyDeviceToHost, s));
    if (arena_used_) VR_TRY_HIP(hipMemcpyAsync(arena.data(), arena_.p, (size_t)arena_used_, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    offsets->assign(1, 0);
    bytes->clear();
    bytes->reserve((size_t)arena_used_);
    for (int64_t id = 0; id < count_; id++) {
        const uint64_t sp = spans[(size_t)id];
        const size_t off = (size_t)(sp >> 32), len = (size_t)(uint32_t)sp;
        bytes->insert(bytes->end(), arena.begin() + off, arena.begin() + off + len);
        offsets->push_back((int32_t)bytes->size());
    }
    return VR_OK;
}

}  // namespace vr
