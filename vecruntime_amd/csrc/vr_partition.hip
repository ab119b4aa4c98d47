// Shuffle partitioning on gfx950: vr_hash_partition_ids, vr_round_robin_ids and vr_partition_rows
// of include/vr.h.
//
// Replaces the map side's partition-id step and row grouping of the columnar shuffle
// (VectorShuffleManager.partitionIds, shuffle/src/main/scala/org/apache/spark/sql/vecruntime/shuffle/
// VectorShuffleManager.scala:332-350; PartitionedIpcWriter, shuffle/src/main/scala/io/vecruntime/
// shuffle/PartitionedIpcWriter.scala:1003,1150-1154) with the rules of PartitionKernels
// (kernels/src/main/java/io/vecruntime/kernels/PartitionKernels.java): Spark's
// Pmod(Murmur3Hash(keys, 42), n), bit-identical, so Spark's co-partitioning and exchange reuse hold.
//
// The hash is a streaming pass (one thread per row, every key mixed in registers, one int32 id
// written): HBM-bound at the keys' bytes plus four or eight per row.  The grouping up to
// MS_MAX_PARTS partitions is a two-pass counting sort: per-tile partition counts, an exclusive scan
// of them partition-major, then each row's slot -- written in row order (dest) and as the row at
// its slot (rows).  Above that it is a stable radix sort of (id, row) over the bits an id needs.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "vr_device.h"
#include "vr_internal.h"

namespace vr {
namespace {

struct HashKey {
    const void* data;
    const uint8_t* valid;
    const int32_t* offsets;    // plain UTF8
    const int32_t* dict_off;   // dictionary-encoded UTF8: the dictionary's offsets and bytes
    const uint8_t* dict_data;
    int32_t kind;              // vr_hash_kind
    int32_t dict_len;
};

// The key descriptors live in device memory and are read with uniform (scalar) loads: an array of
// them passed by value and indexed by the key number would be copied to every lane's scratch.
struct HashArgs {
    const HashKey* __restrict__ keys;
    int32_t nkeys;
    int32_t parts;
    uint32_t seed;
    int64_t n;
    int32_t* ids;
    uint32_t* hashes;          // may be NULL
    uint32_t* err;             // set to 1 for a dictionary code outside its dictionary
};

// Spark's Murmur3_x86_32 (org.apache.spark.unsafe.hash.Murmur3_x86_32, restated in
// PartitionKernels.java:89-163): 32-bit words mixed into h1, then fmix with the byte length.
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ uint32_t mix_k1(uint32_t k1) {
    k1 *= 0xcc9e2d51u;
    k1 = rotl32(k1, 15);
    return k1 * 0x1b873593u;
}

__device__ __forceinline__ uint32_t mix_h1(uint32_t h1, uint32_t k1) {
    h1 ^= k1;
    h1 = rotl32(h1, 13);
    return h1 * 5u + 0xe6546b64u;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t h1, uint32_t len) {
    h1 ^= len;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    h1 ^= h1 >> 16;
    return h1;
}

__device__ __forceinline__ uint32_t hash_int(uint32_t v, uint32_t seed) { return fmix32(mix_h1(seed, mix_k1(v)), 4); }

__device__ __forceinline__ uint32_t hash_long(uint64_t v, uint32_t seed) {
    const uint32_t h1 = mix_h1(seed, mix_k1((uint32_t)v));
    return fmix32(mix_h1(h1, mix_k1((uint32_t)(v >> 32))), 8);
}

// hashUnsafeBytes: little-endian 4-byte words, then each remaining byte sign-extended and mixed as
// a word of its own (Spark's tail, not the standard Murmur3 one).  Strings start at any byte, so a
// word is assembled from bytes.
__device__ __forceinline__ uint32_t hash_bytes(const uint8_t* p, int32_t len, uint32_t seed) {
    uint32_t h1 = seed;
    const int32_t aligned = len & ~3;
    for (int32_t i = 0; i < aligned; i += 4) {
        const uint32_t w = (uint32_t)p[i] | ((uint32_t)p[i + 1] << 8) | ((uint32_t)p[i + 2] << 16) |
                           ((uint32_t)p[i + 3] << 24);
        h1 = mix_h1(h1, mix_k1(w));
    }
    for (int32_t i = aligned; i < len; i++) h1 = mix_h1(h1, mix_k1((uint32_t)(int32_t)(int8_t)p[i]));
    return fmix32(h1, (uint32_t)len);
}

// A decimal above 18 digits: hashUnsafeBytes over BigInteger.toByteArray of the unscaled value,
// the minimal big-endian two's-complement form (bitLength / 8 + 1 bytes, 1 to 16).
__device__ __forceinline__ uint32_t hash_decimal(uint64_t lo, int64_t hi, uint32_t seed) {
    uint64_t ml = lo, mh = (uint64_t)hi;
    if (hi < 0) {
        ml = ~ml;
        mh = ~mh;
    }
    const int bitlen = mh ? 128 - __clzll((long long)mh) : (ml ? 64 - __clzll((long long)ml) : 0);
    const int nb = bitlen / 8 + 1;
    auto byte_at = [&](int j) -> uint32_t {  // byte j of nb, most significant first
        const int sh = 8 * (nb - 1 - j);
        const uint64_t w = sh >= 64 ? ((uint64_t)hi >> (sh - 64)) : sh == 0 ? lo : ((lo >> sh) | ((uint64_t)hi << (64 - sh)));
        return (uint32_t)(w & 0xff);
    };
    uint32_t h1 = seed;
    const int aligned = nb & ~3;
    for (int j = 0; j < aligned; j += 4)
        h1 = mix_h1(h1, mix_k1(byte_at(j) | (byte_at(j + 1) << 8) | (byte_at(j + 2) << 16) | (byte_at(j + 3) << 24)));
    for (int j = aligned; j < nb; j++) h1 = mix_h1(h1, mix_k1((uint32_t)(int32_t)(int8_t)byte_at(j)));
    return fmix32(h1, (uint32_t)nb);
}

// Spark's normalize before hashing: both zeros hash as 0.0, every NaN as the canonical one
// (PartitionKernels.doubleBits / floatBits).
__device__ __forceinline__ uint64_t double_bits(double v) {
    if (v == 0.0) return 0;
    if (v != v) return 0x7ff8000000000000ull;
    return (uint64_t)__double_as_longlong(v);
}

__device__ __forceinline__ uint32_t float_bits(float v) {
    if (v == 0.0f) return 0;
    if (v != v) return 0x7fc00000u;
    return (uint32_t)__float_as_int(v);
}

__device__ __forceinline__ bool valid_at(const uint8_t* bm, int64_t i) { return !bm || ((bm[i >> 3] >> (i & 7)) & 1u); }

__device__ __forceinline__ uint32_t mix_key(const HashKey& k, int64_t i, uint32_t h, uint32_t* err) {
    switch (k.kind) {
        case VR_HASH_INT: return hash_int((uint32_t)((const int32_t*)k.data)[i], h);
        case VR_HASH_LONG: return hash_long((uint64_t)((const int64_t*)k.data)[i], h);
        case VR_HASH_DOUBLE: return hash_long(double_bits(((const double*)k.data)[i]), h);
        case VR_HASH_FLOAT: return hash_int(float_bits((float)((const double*)k.data)[i]), h);
        case VR_HASH_BOOL: return hash_int((((const uint8_t*)k.data)[i >> 3] >> (i & 7)) & 1u, h);
        case VR_HASH_DECIMAL128: {
            const uint64_t* d = (const uint64_t*)k.data;
            return hash_decimal(d[2 * i], (int64_t)d[2 * i + 1], h);
        }
        default: {  // VR_HASH_UTF8
            if (k.dict_off) {
                const int32_t code = ((const int32_t*)k.data)[i];
                if (code < 0 || code >= k.dict_len) {
                    *err = 1;
                    return h;
                }
                const int32_t a = k.dict_off[code];
                return hash_bytes(k.dict_data + a, k.dict_off[code + 1] - a, h);
            }
            const int32_t a = k.offsets[i];
            return hash_bytes((const uint8_t*)k.data + a, k.offsets[i + 1] - a, h);
        }
    }
}

constexpr int HASH_ROWS = 4;  // rows per thread per step, their loads issued before any is hashed

// One key mixed into HASH_ROWS rows' running hashes: the kind is uniform, so the switch sits outside
// the rows and every fixed-width value is loaded before the first is hashed.
__device__ __forceinline__ void mix_rows(const HashKey& k, const int64_t (&i)[HASH_ROWS], const bool (&in)[HASH_ROWS],
                                         uint32_t (&h)[HASH_ROWS], uint32_t* err) {
    bool live[HASH_ROWS];
#pragma unroll
    for (int u = 0; u < HASH_ROWS; u++) live[u] = in[u] && valid_at(k.valid, i[u]);
    switch (k.kind) {
        case VR_HASH_INT: {
            uint32_t x[HASH_ROWS];
#pragma unroll
            for (int u = 0; u < HASH_ROWS; u++) x[u] = in[u] ? (uint32_t)((const int32_t*)k.data)[i[u]] : 0;
#pragma unroll
            for (int u = 0; u < HASH_ROWS; u++) if (live[u]) h[u] = hash_int(x[u], h[u]);
            break;
        }
        case VR_HASH_LONG:
        case VR_HASH_DOUBLE:
        case VR_HASH_FLOAT: {
            uint64_t x[HASH_ROWS];
#pragma unroll
            for (int u = 0; u < HASH_ROWS; u++) x[u] = in[u] ? ((const uint64_t*)k.data)[i[u]] : 0;
#pragma unroll
            for (int u = 0; u < HASH_ROWS; u++) {
                if (!live[u]) continue;
                if (k.kind == VR_HASH_LONG) h[u] = hash_long(x[u], h[u]);
                else if (k.kind == VR_HASH_DOUBLE) h[u] = hash_long(double_bits(__longlong_as_double((long long)x[u])), h[u]);
                else h[u] = hash_int(float_bits((float)__longlong_as_double((long long)x[u])), h[u]);
            }
            break;
        }
        default:
#pragma unroll
            for (int u = 0; u < HASH_ROWS; u++) if (live[u]) h[u] = mix_key(k, i[u], h[u], err);
    }
}

__global__ void __launch_bounds__(256) hash_ids_kernel(HashArgs a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < a.n; i0 += HASH_ROWS * stride) {
        int64_t i[HASH_ROWS];
        bool in[HASH_ROWS];
        uint32_t h[HASH_ROWS];
#pragma unroll
        for (int u = 0; u < HASH_ROWS; u++) {
            i[u] = i0 + u * stride;
            in[u] = i[u] < a.n;
            h[u] = a.seed;
        }
        for (int c = 0; c < a.nkeys; c++) mix_rows(a.keys[c], i, in, h, a.err);  // a null key leaves the hash
#pragma unroll
        for (int u = 0; u < HASH_ROWS; u++) {
            if (!in[u]) continue;
            if (a.hashes) a.hashes[i[u]] = h[u];
            const int32_t r = (int32_t)h[u] % a.parts;  // Pmod: the non-negative remainder
            a.ids[i[u]] = r < 0 ? r + a.parts : r;
        }
    }
}

__global__ void __launch_bounds__(256) round_robin_kernel(int32_t* ids, int64_t n, int32_t parts, int32_t start) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        ids[i] = (int32_t)(((int64_t)start + i) % parts);
}

__global__ void __launch_bounds__(256) iota_kernel(int32_t* v, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (int32_t)i;
}

// starts[p] = the first slot of partition p in the sorted ids: slot k opens every partition in
// (sorted[k - 1], sorted[k]], the last slot closes the ones after it with n.  An id outside
// [0, parts) flags `err`; the writes stay clamped to starts[0 .. parts].
__global__ void __launch_bounds__(256) starts_kernel(const uint32_t* sorted, int64_t n, int32_t parts, int64_t* starts,
                                                     uint32_t* err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const uint32_t cur = sorted[k];
        if (cur >= (uint32_t)parts) *err = 1;
        const int64_t c = std::min<int64_t>(cur, parts - 1);
        const int64_t prev = k ? std::min<int64_t>(sorted[k - 1], parts - 1) : -1;
        for (int64_t p = prev + 1; p <= c; p++) starts[p] = k;
        if (k == n - 1)
            for (int64_t p = c + 1; p <= parts; p++) starts[p] = n;
    }
}

__global__ void __launch_bounds__(256) fill_i64_kernel(int64_t* v, int64_t count, int64_t value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) v[i] = value;
}

__global__ void __launch_bounds__(256) inverse_kernel(const int32_t* rows, int64_t n, int32_t* dest) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) dest[rows[k]] = (int32_t)k;
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0; }

// The counting sort for up to MS_MAX_PARTS partitions.  A block takes a tile of MS_TILE rows, each
// of its four waves MS_WAVE_ROWS consecutive rows of it in chunks of 64.  Within a chunk the lanes
// holding one id find each other with one ballot per id bit (the warp-level multisplit of Ashkiani
// et al., "GPU Multisplit", PPoPP 2016); the lowest of them adds the group's size to the wave's
// LDS count for the id, and a lane's rank is the number of its peers below it.  Chunks in order,
// waves in order, tiles in order: the grouping keeps row order within a partition.
constexpr int MS_ROWS = 16;                   // rows per lane
constexpr int MS_WAVE_ROWS = 64 * MS_ROWS;     // consecutive rows per wave
constexpr int MS_TILE = 4 * MS_WAVE_ROWS;      // rows per block of 256 threads
constexpr int MS_MAX_PARTS = 512;

// The tile a block takes: blocks are dispatched to the 8 XCDs in turn, so XCD x gets a contiguous
// run of tiles and its blocks' counts (adjacent words for one partition) meet in its own L2.  A
// bijection of [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_tile() {
    const uint32_t xcds = 8, g = gridDim.x, b = blockIdx.x;
    const uint32_t q = g / xcds, r = g % xcds, x = b % xcds, k = b / xcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__device__ __forceinline__ uint64_t peer_lanes(uint32_t id, int bits, uint64_t live) {
    uint64_t m = live;
    for (int b = 0; b < bits; b++) {
        const bool on = (id >> b) & 1u;
        const uint64_t set = __ballot(on);
        m &= on ? set : ~set;
    }
    return m;
}

// Loads this lane's MS_ROWS ids; a row past n or with an id outside [0, parts) is not live (the
// count pass flags the latter).
__device__ __forceinline__ void ms_load(const int32_t* ids, int64_t base, int64_t n, int32_t parts, uint32_t (&v)[MS_ROWS],
                                        uint32_t& okmask) {
    const int lane = threadIdx.x & 63;
    okmask = 0;
#pragma unroll
    for (int c = 0; c < MS_ROWS; c++) {
        const int64_t i = base + c * 64 + lane;
        const uint32_t id = i < n ? (uint32_t)ids[i] : 0xffffffffu;
        const bool ok = id < (uint32_t)parts;
        v[c] = ok ? id : 0;
        okmask |= (uint32_t)ok << c;
    }
}

__global__ void __launch_bounds__(256) ms_count_kernel(const int32_t* ids, int64_t n, int32_t parts, int bits,
                                                       int64_t tiles, int32_t* counts, uint32_t* err) {
    __shared__ int32_t hist[MS_MAX_PARTS];
    for (int p = threadIdx.x; p < parts; p += 256) hist[p] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint32_t tile = xcd_tile();
    const int64_t base = (int64_t)tile * MS_TILE + (int64_t)(threadIdx.x >> 6) * MS_WAVE_ROWS;
    uint32_t v[MS_ROWS], okmask;
    ms_load(ids, base, n, parts, v, okmask);
    bool bad = false;
    for (int c = 0; c < MS_ROWS; c++) {
        const bool ok = (okmask >> c) & 1u;
        bad |= !ok && base + c * 64 + lane < n;
        const uint64_t m = peer_lanes(v[c], bits, __ballot(ok));
        if (ok && lane == __builtin_ctzll(m)) atomicAdd(&hist[v[c]], __popcll(m));
    }
    if (bad) *err = 1;
    __syncthreads();
    for (int p = threadIdx.x; p < parts; p += 256) counts[(int64_t)p * tiles + tile] = hist[p];
}

// offsets: the exclusive scan of counts (partition-major), the first slot of partition p's rows of
// tile b.  A tile's rows of one partition take consecutive slots, so each row is first placed in
// LDS at its slot less the partition's delta (its first slot less the rows of lower partitions in
// the tile) and the tile's `rows` then leave in runs of one partition each.
__global__ void __launch_bounds__(256) ms_place_kernel(const int32_t* ids, int64_t n, int32_t parts, int bits,
                                                       int64_t tiles, const int32_t* offsets, int32_t* rows,
                                                       int32_t* dest) {
    typedef hipcub::BlockScan<int32_t, 256> TileScan;
    __shared__ int32_t next[4 * MS_MAX_PARTS];  // per wave: the next free slot of each partition
    __shared__ int32_t delta[MS_MAX_PARTS];
    __shared__ int32_t srow[MS_TILE], sslot[MS_TILE];
    __shared__ typename TileScan::TempStorage scan_tmp;
    __shared__ int32_t s_live;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t tile = xcd_tile();
    const int64_t base = (int64_t)tile * MS_TILE + (int64_t)w * MS_WAVE_ROWS;
    uint32_t v[MS_ROWS], okmask;
    ms_load(ids, base, n, parts, v, okmask);
    for (int p = threadIdx.x; p < 4 * parts; p += 256) next[p] = 0;
    __syncthreads();
    uint64_t peers[MS_ROWS];  // kept for the placing below
#pragma unroll
    for (int c = 0; c < MS_ROWS; c++) {  // this wave's counts
        const bool ok = (okmask >> c) & 1u;
        peers[c] = peer_lanes(v[c], bits, __ballot(ok));
        if (ok && lane == __builtin_ctzll(peers[c])) atomicAdd(&next[w * parts + v[c]], __popcll(peers[c]));
    }
    __syncthreads();
    // partitions 2t and 2t + 1: the tile's first slot, then the waves in order; the tile's total
    int32_t first[2], total[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int p = 2 * threadIdx.x + k;
        first[k] = total[k] = 0;
        if (p < parts) {
            int32_t run = first[k] = offsets[(int64_t)p * tiles + tile];
            for (int q = 0; q < 4; q++) {
                const int32_t t = next[q * parts + p];
                next[q * parts + p] = run;
                run += t;
            }
            total[k] = run - first[k];
        }
    }
    int32_t live;
    TileScan(scan_tmp).ExclusiveSum(total, total, live);
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (2 * (int)threadIdx.x + k < parts) delta[2 * threadIdx.x + k] = first[k] - total[k];
    if (threadIdx.x == 0) s_live = live;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MS_ROWS; c++) {
        const bool ok = (okmask >> c) & 1u;
        const uint64_t m = peers[c];
        const int leader = ok ? __builtin_ctzll(m) : lane;
        int32_t at = 0;
        if (ok && lane == leader) at = atomicAdd(&next[w * parts + v[c]], __popcll(m));
        at = __shfl(at, leader);
        if (ok) {
            const int64_t i = base + c * 64 + lane;
            const int32_t slot = at + __popcll(m & ((1ull << lane) - 1));
            if (dest) dest[i] = slot;
            const int32_t at_tile = slot - delta[v[c]];
            srow[at_tile] = (int32_t)i;
            sslot[at_tile] = slot;
        }
    }
    __syncthreads();
    if (rows)
        for (int j = threadIdx.x; j < s_live; j += 256) rows[sslot[j]] = srow[j];
}

__global__ void __launch_bounds__(256) ms_starts_kernel(const int32_t* offsets, int32_t parts, int64_t tiles, int64_t n,
                                                        int64_t* starts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= parts; p += stride)
        starts[p] = p < parts ? offsets[p * tiles] : n;
}

int grid_for(uint64_t work, int cap) {
    const uint64_t g = (work + 255) / 256;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)cap));
}

int sort_bits(int32_t parts) {  // bits that hold every id in [0, parts)
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < parts) b++;
    return b;
}

}  // namespace
}  // namespace vr

using namespace vr;

extern "C" {

vr_status vr_hash_partition_ids(vr_device* dev, const vr_column* keys, const int32_t* kinds, int32_t num_keys,
                                int64_t n, int32_t num_partitions, int32_t seed, int32_t* ids, int32_t* hashes,
                                void* stream) {
    if (!dev || (num_keys && (!keys || !kinds)) || (!ids && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (num_keys < 0 || n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative count");
    if (num_partitions < 1) return fail(VR_ERR_INVALID_ARGUMENT, "partition count below 1");
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    std::vector<HashKey> hk((size_t)num_keys);
    std::vector<DevBuf> dicts;
    for (int c = 0; c < num_keys; c++) {
        const vr_column& k = keys[c];
        if (k.mem != VR_MEM_DEVICE) return fail(VR_ERR_UNSUPPORTED, "partition keys must be device-resident");
        if (k.length < n) return fail(VR_ERR_INVALID_ARGUMENT, "key column shorter than the batch");
        if (!k.data && n) return fail(VR_ERR_INVALID_ARGUMENT, "key column without data");
        int32_t lane;
        switch (kinds[c]) {
            case VR_HASH_INT: lane = VR_INT32; break;
            case VR_HASH_LONG: lane = VR_INT64; break;
            case VR_HASH_DOUBLE:
            case VR_HASH_FLOAT: lane = VR_FLOAT64; break;
            case VR_HASH_BOOL: lane = VR_BOOL; break;
            case VR_HASH_UTF8: lane = VR_UTF8; break;
            case VR_HASH_DECIMAL128: lane = VR_DECIMAL128; break;
            default: return fail(VR_ERR_INVALID_ARGUMENT, "unknown hash kind");
        }
        if (k.type != lane) return fail(VR_ERR_INVALID_ARGUMENT, "hash kind does not fit the key's lane");
        HashKey& h = hk[(size_t)c];
        h = HashKey{k.data, k.validity, k.offsets, nullptr, nullptr, kinds[c], 0};
        if (lane == VR_UTF8 && k.dictionary) {
            const vr_column& dc = *k.dictionary;
            if (dc.type != VR_UTF8 || dc.dictionary || !dc.offsets) return fail(VR_ERR_INVALID_ARGUMENT, "a dictionary is a plain UTF8 column");
            h.dict_len = (int32_t)dc.length;
            if (dc.mem == VR_MEM_DEVICE) {
                h.dict_off = dc.offsets;
                h.dict_data = (const uint8_t*)dc.data;
            }
        } else if (lane == VR_UTF8 && !k.offsets) {
            return fail(VR_ERR_INVALID_ARGUMENT, "a plain UTF8 key needs offsets");
        }
    }
    if (n == 0) return VR_OK;
    // every key checked: host dictionaries are copied over for this call
    for (int c = 0; c < num_keys; c++) {
        const vr_column& k = keys[c];
        if (k.type != VR_UTF8 || !k.dictionary || k.dictionary->mem == VR_MEM_DEVICE) continue;
        const vr_column& dc = *k.dictionary;
        const int32_t nbytes = dc.length ? dc.offsets[dc.length] : 0;
        DevBuf off, bytes;
        VR_TRY(off.ensure(&d, ((size_t)dc.length + 1) * 4));
        VR_TRY(bytes.ensure(&d, (size_t)std::max(nbytes, 1)));
        VR_TRY_HIP(hipMemcpyAsync(off.p, dc.offsets, ((size_t)dc.length + 1) * 4, hipMemcpyHostToDevice, s));
        if (nbytes) VR_TRY_HIP(hipMemcpyAsync(bytes.p, dc.data, (size_t)nbytes, hipMemcpyHostToDevice, s));
        hk[(size_t)c].dict_off = off.as<int32_t>();
        hk[(size_t)c].dict_data = bytes.as<uint8_t>();
        dicts.push_back(std::move(off));
        dicts.push_back(std::move(bytes));
    }
    DevBuf err, kbuf;
    VR_TRY(err.ensure(&d, 4));
    VR_TRY(kbuf.ensure(&d, sizeof(HashKey) * (size_t)std::max(1, num_keys)));
    if (num_keys) VR_TRY_HIP(hipMemcpyAsync(kbuf.p, hk.data(), sizeof(HashKey) * (size_t)num_keys, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, err.as<uint32_t>());
    HashArgs a{kbuf.as<HashKey>(), num_keys, num_partitions, (uint32_t)seed, n, ids, (uint32_t*)hashes, err.as<uint32_t>()};
    hipLaunchKernelGGL(hash_ids_kernel, dim3(grid_for(((uint64_t)n + HASH_ROWS - 1) / HASH_ROWS, d.num_cus * 32)), dim3(256), 0,
                       s, a);
    VR_TRY_HIP(hipGetLastError());
    uint32_t bad = 0;
    VR_TRY_HIP(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (bad) return fail(VR_ERR_INVALID_ARGUMENT, "a dictionary code outside its dictionary");
    d.count(&vr_stats::calls, 1);
    d.count(&vr_stats::input_rows, n);
    d.count(&vr_stats::kernel_launches, 2);
    return VR_OK;
}

vr_status vr_round_robin_ids(vr_device* dev, int64_t n, int32_t num_partitions, int32_t start, int32_t* ids,
                             int32_t* next, void* stream) {
    if (!dev || (!ids && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative count");
    if (num_partitions < 1) return fail(VR_ERR_INVALID_ARGUMENT, "partition count below 1");
    if (start < 0 || start >= num_partitions) return fail(VR_ERR_INVALID_ARGUMENT, "start outside the partitions");
    if (next) *next = (int32_t)(((int64_t)start + n) % num_partitions);
    if (n == 0) return VR_OK;
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    hipLaunchKernelGGL(round_robin_kernel, dim3(grid_for((uint64_t)n, d.num_cus * 32)), dim3(256), 0, s, ids, n,
                       num_partitions, start);
    VR_TRY_HIP(hipGetLastError());
    VR_TRY_HIP(hipStreamSynchronize(s));
    d.count(&vr_stats::calls, 1);
    return VR_OK;
}

vr_status vr_partition_rows(vr_device* dev, const int32_t* ids, int64_t n, int32_t num_partitions, int32_t* rows,
                            int32_t* dest, int64_t* starts, void* stream) {
    if (!dev || !starts || (!ids && n)) return fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    if (n < 0) return fail(VR_ERR_INVALID_ARGUMENT, "negative count");
    if (n >= (1ll << 31)) return fail(VR_ERR_UNSUPPORTED, "partitioning of batches past 2^31 rows");
    if (num_partitions < 1) return fail(VR_ERR_INVALID_ARGUMENT, "partition count below 1");
    Device& d = dev->d;
    hipStream_t s = d.pick(stream);
    const int cap = d.num_cus * 32;
    if (n == 0) {
        hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for((uint64_t)num_partitions + 1, cap)), dim3(256), 0, s, starts,
                           (int64_t)num_partitions + 1, (int64_t)0);
        VR_TRY_HIP(hipGetLastError());
        VR_TRY_HIP(hipStreamSynchronize(s));
        return VR_OK;
    }
    const int bits = sort_bits(num_partitions);
    if (num_partitions <= MS_MAX_PARTS) {
        const int64_t tiles = (n + MS_TILE - 1) / MS_TILE;
        const int64_t cells = tiles * num_partitions;
        DevBuf counts, offsets, tmp, err;
        VR_TRY(counts.ensure(&d, (size_t)cells * 4));
        VR_TRY(offsets.ensure(&d, (size_t)cells * 4));
        VR_TRY(err.ensure(&d, 4));
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, err.as<uint32_t>());
        hipLaunchKernelGGL(ms_count_kernel, dim3((unsigned)tiles), dim3(256), 0, s, ids, n, num_partitions, bits, tiles,
                           counts.as<int32_t>(), err.as<uint32_t>());
        size_t tb = 0;
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, counts.as<int32_t>(), offsets.as<int32_t>(), (int)cells, s));
        VR_TRY(tmp.ensure(&d, std::max<size_t>(tb, 1)));
        VR_TRY_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, counts.as<int32_t>(), offsets.as<int32_t>(), (int)cells, s));
        if (rows || dest)
            hipLaunchKernelGGL(ms_place_kernel, dim3((unsigned)tiles), dim3(256), 0, s, ids, n, num_partitions, bits, tiles,
                               offsets.as<int32_t>(), rows, dest);
        hipLaunchKernelGGL(ms_starts_kernel, dim3(grid_for((uint64_t)num_partitions + 1, cap)), dim3(256), 0, s,
                           offsets.as<int32_t>(), num_partitions, tiles, n, starts);
        VR_TRY_HIP(hipGetLastError());
        uint32_t bad = 0;
        VR_TRY_HIP(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, s));
        VR_TRY_HIP(hipStreamSynchronize(s));
        if (bad) return fail(VR_ERR_INVALID_ARGUMENT, "a partition id outside [0, num_partitions)");
        d.count(&vr_stats::calls, 1);
        d.count(&vr_stats::input_rows, n);
        return VR_OK;
    }
    DevBuf sorted, iota, own_rows, tmp, err;
    VR_TRY(sorted.ensure(&d, (size_t)n * 4));
    VR_TRY(iota.ensure(&d, (size_t)n * 4));
    VR_TRY(err.ensure(&d, 4));
    int32_t* order = rows;
    if (!order) {
        VR_TRY(own_rows.ensure(&d, (size_t)n * 4));
        order = own_rows.as<int32_t>();
    }
    const int grid = grid_for((uint64_t)n, cap);
    hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, s, iota.as<int32_t>(), n);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, err.as<uint32_t>());
    size_t tb = 0;
    VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const uint32_t*)ids, sorted.as<uint32_t>(),
                                                  iota.as<int32_t>(), order, (int)n, 0, bits, s));
    VR_TRY(tmp.ensure(&d, std::max<size_t>(tb, 1)));
    VR_TRY_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, (const uint32_t*)ids, sorted.as<uint32_t>(),
                                                  iota.as<int32_t>(), order, (int)n, 0, bits, s));
    hipLaunchKernelGGL(starts_kernel, dim3(grid), dim3(256), 0, s, sorted.as<uint32_t>(), n, num_partitions, starts,
                       err.as<uint32_t>());
    if (dest) hipLaunchKernelGGL(inverse_kernel, dim3(grid), dim3(256), 0, s, order, n, dest);
    VR_TRY_HIP(hipGetLastError());
    uint32_t bad = 0;
    VR_TRY_HIP(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, s));
    VR_TRY_HIP(hipStreamSynchronize(s));
    if (bad) return fail(VR_ERR_INVALID_ARGUMENT, "a partition id outside [0, num_partitions)");
    d.count(&vr_stats::calls, 1);
    d.count(&vr_stats::input_rows, n);
    return VR_OK;
}

}  // extern "C"
