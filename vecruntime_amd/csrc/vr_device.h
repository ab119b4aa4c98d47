// Device-side helpers shared by the kernels: bitmaps, hashing, ordered encodings, exact
// multi-limb atomics and 128/192-bit decimal arithmetic.
#pragma once

#include "vr_types.h"

namespace vr {

// The empty marker of a 64-bit key cell.  A real key equal to it is kept in a reserved row.
constexpr uint64_t EMPTY_KEY = 0x7FF4DEADBEEF7FF4ull;

// Pointers into device (or host-mapped) memory, never LDS: loads through them may be reordered
// with LDS atomics.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}

// Bit i of an Arrow bitmap (validity or selection) in device memory.
__device__ __forceinline__ bool bit_at(const uint8_t* bm, int64_t i) { return (gptr(bm)[i >> 3] >> (i & 7)) & 1; }

__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Waits for every vector memory operation of this wave (loads and stores) to complete: the s_waitcnt
// encoding of gfx9 with vmcnt 0 and expcnt / lgkmcnt left at their maxima (no wait on them).
__device__ __forceinline__ void vmem_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Streaming loads of input columns: read once, so they should not displace the hash table from
// the caches (nontemporal hint).
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(gptr(p));
}

// Coherent (agent-scope) load, for cells another workgroup may have written during this kernel.
__device__ __forceinline__ uint64_t ld_coherent(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t atomic_add_u64(uint64_t* p, uint64_t v) {
    return atomicAdd((unsigned long long*)p, (unsigned long long)v);
}
__device__ __forceinline__ uint64_t atomic_cas_u64(uint64_t* p, uint64_t expect, uint64_t v) {
    return atomicCAS((unsigned long long*)p, (unsigned long long)expect, (unsigned long long)v);
}
__device__ __forceinline__ void atomic_min_u64(uint64_t* p, uint64_t v) { atomicMin((unsigned long long*)p, (unsigned long long)v); }
__device__ __forceinline__ void atomic_max_u64(uint64_t* p, uint64_t v) { atomicMax((unsigned long long*)p, (unsigned long long)v); }
__device__ __forceinline__ void atomic_min_i64(uint64_t* p, int64_t v) { atomicMin((long long*)p, (long long)v); }
__device__ __forceinline__ void atomic_max_i64(uint64_t* p, int64_t v) { atomicMax((long long*)p, (long long)v); }
__device__ __forceinline__ void atomic_add_f64(uint64_t* p, double v) { unsafeAtomicAdd((double*)p, v); }

// (hi, lo) += (xh, xl) as a signed 128-bit value.  The low limb's atomic returns the running
// value it was added to, so its carry is exact whatever other rows interleave; the high limb then
// takes xh + carry.  Exact for any order of the rows.
__device__ __forceinline__ void atomic_add_i128(uint64_t* lo, uint64_t* hi, uint64_t xl, uint64_t xh) {
    uint64_t old = atomic_add_u64(lo, xl);
    uint64_t carry = (old + xl) < old ? 1ull : 0ull;
    uint64_t t = xh + carry;
    if (t) atomic_add_u64(hi, t);
}

// (s, hi, lo) += sign-extended (xh, xl): a 192-bit exact sum of 128-bit values, whose third limb
// tells whether the total left the signed 128-bit range.
__device__ __forceinline__ void atomic_add_i192(uint64_t* lo, uint64_t* hi, uint64_t* s, uint64_t xl, uint64_t xh) {
    uint64_t old = atomic_add_u64(lo, xl);
    uint64_t c0 = (old + xl) < old ? 1ull : 0ull;
    uint64_t t = xh + c0;
    uint64_t tc = (t < xh) ? 1ull : 0ull;  // xh = ~0 with a carry in wraps to 0 and carries out
    uint64_t c1 = tc;
    if (t) {
        uint64_t oh = atomic_add_u64(hi, t);
        c1 += (oh + t) < oh ? 1ull : 0ull;
    }
    uint64_t xs = ((int64_t)xh < 0) ? ~0ull : 0ull;
    uint64_t ts = xs + c1;
    if (ts) atomic_add_u64(s, ts);
}

// Order-preserving unsigned encoding of a double for atomic min/max: NaN (any payload) maps to
// the greatest code and -0.0 to +0.0's code, as CompareOp.nanSafeCompare orders them
// (CompareOp.java:68-82).  The payload and the zero's sign are recovered separately.
__host__ __device__ __forceinline__ uint64_t order_f64(uint64_t bits) {
    if ((bits & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) bits = 0x7FF8000000000000ull;
    if (bits == 0x8000000000000000ull) bits = 0;
    return (bits & 0x8000000000000000ull) ? ~bits : (bits | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ uint64_t unorder_f64(uint64_t o) {
    return (o & 0x8000000000000000ull) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
}
__host__ __device__ __forceinline__ bool is_nan_bits(uint64_t b) { return (b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull; }
__host__ __device__ __forceinline__ bool is_zero_bits(uint64_t b) { return (b & 0x7FFFFFFFFFFFFFFFull) == 0; }

// ------------------------------------------------------------------ decimal arithmetic

// 10^k as (hi, lo) for k = 0..38.
struct U128 {
    uint64_t lo, hi;
};
__device__ __forceinline__ U128 pow10_128(int k) {
    U128 r{1, 0};
    for (int i = 0; i < k; i++) {
        // r *= 10
        uint64_t lo_lo = (r.lo & 0xFFFFFFFFull) * 10ull;
        uint64_t lo_hi = (r.lo >> 32) * 10ull + (lo_lo >> 32);
        uint64_t nlo = (lo_hi << 32) | (lo_lo & 0xFFFFFFFFull);
        uint64_t carry = lo_hi >> 32;
        r.hi = r.hi * 10ull + carry;
        r.lo = nlo;
    }
    return r;
}

// Magnitude of a signed 128-bit value.
__device__ __forceinline__ U128 abs128(uint64_t lo, uint64_t hi, bool* neg) {
    *neg = (int64_t)hi < 0;
    if (!*neg) return U128{lo, hi};
    uint64_t nlo = ~lo + 1;
    uint64_t nhi = ~hi + (nlo == 0 ? 1ull : 0ull);
    return U128{nlo, nhi};
}
__device__ __forceinline__ bool ge128(U128 a, U128 b) { return a.hi != b.hi ? a.hi > b.hi : a.lo >= b.lo; }

// |v| >= 10^p for a signed 128-bit v (p <= 38).
__device__ __forceinline__ bool past_precision(uint64_t lo, uint64_t hi, int p) {
    if (p >= 39) return false;
    bool neg;
    U128 m = abs128(lo, hi, &neg);
    return ge128(m, pow10_128(p));
}

// (r:hi, lo) / d with hi < d: quotient (64 bits) and remainder, by shift and subtract.
__device__ __forceinline__ uint64_t udiv128_64(uint64_t hi, uint64_t lo, uint64_t d, uint64_t* rem) {
    uint64_t q = 0;
    uint64_t r = hi;
    for (int i = 63; i >= 0; i--) {
        uint64_t top = r >> 63;
        r = (r << 1) | ((lo >> i) & 1ull);
        q <<= 1;
        if (top || r >= d) {
            r -= d;
            q |= 1ull;
        }
    }
    *rem = r;
    return q;
}

// Spark's decimal average: sum (scale s) / count at scale rs, rounded half away from zero
// (DecimalDivideWithOverflowCheck's HALF_UP).  Returns false when the quotient leaves 127 bits.
__device__ __forceinline__ bool decimal_avg(uint64_t lo, uint64_t hi, uint64_t count, int k, uint64_t* qlo,
                                            uint64_t* qhi) {
    bool neg;
    U128 m = abs128(lo, hi, &neg);
    uint64_t w[3] = {m.lo, m.hi, 0};
    for (int t = 0; t < k; t++) {  // w *= 10 over three limbs
        uint64_t carry = 0;
        for (int j = 0; j < 3; j++) {
            uint64_t x = w[j];
            uint64_t plo = (x & 0xFFFFFFFFull) * 10ull + carry;
            uint64_t phi = (x >> 32) * 10ull + (plo >> 32);
            w[j] = (phi << 32) | (plo & 0xFFFFFFFFull);
            carry = phi >> 32;
        }
    }
    uint64_t r = 0, q[3];
    for (int j = 2; j >= 0; j--) q[j] = udiv128_64(r, w[j], count, &r);
    // half up on the magnitude: 2 * rem >= count (rem < count, so 2*rem may need 65 bits)
    bool up = (r >> 63) || (r << 1) >= count;
    if (up) {
        for (int j = 0; j < 3; j++)
            if (++q[j] != 0) break;
    }
    if (q[2] || (q[1] >> 63)) return false;
    uint64_t rlo = q[0], rhi = q[1];
    if (neg) {
        rlo = ~rlo + 1;
        rhi = ~rhi + (rlo == 0 ? 1ull : 0ull);
    }
    *qlo = rlo;
    *qhi = rhi;
    return true;
}

}  // namespace vr
