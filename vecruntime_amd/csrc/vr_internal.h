// Internal declarations shared by the translation units of libvrgpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "vr_types.h"

namespace vr {

// ------------------------------------------------------------------ errors

vr_status fail(vr_status st, const std::string& msg);
vr_status hip_fail(hipError_t e, const char* what);

#define VR_TRY_HIP(call)                                  \
    do {                                                  \
        hipError_t vr_e_ = (call);                        \
        if (vr_e_ != hipSuccess) return ::vr::hip_fail(vr_e_, #call); \
    } while (0)

#define VR_TRY(call)                      \
    do {                                  \
        vr_status vr_s_ = (call);         \
        if (vr_s_ != VR_OK) return vr_s_; \
    } while (0)

// ------------------------------------------------------------------ device handle

struct Device {
    int ordinal = 0;
    int num_cus = 256;
    int64_t lds_per_block = 64 * 1024;  // LDS bytes a workgroup may use
    hipStream_t stream = nullptr;       // the handle's own compute stream
    hipStream_t copy_stream = nullptr;  // side stream for host <-> device staging
    std::mutex mu;
    int64_t allocated = 0;  // bytes of the blocks handed out
    int64_t budget = 0;
    vr_stats stats{};
    // Open handles (aggregates, joins) plus one for the vr_device_open reference: the device is
    // freed when the last of them lets go, so a handle destroyed after vr_device_close stays valid.
    std::atomic<int64_t> refs{1};
    // Block cache: hipFree waits for the whole device, so blocks no queued work reads any more are
    // kept (by size) and handed out again instead; a request takes the smallest cached block of at
    // least its size and at most 1.5x.  Every block's size is recorded for the release.
    std::multimap<size_t, void*> cache;
    std::unordered_map<void*, size_t> block_bytes;
    int64_t cached = 0;
    int64_t cache_limit = 48ll << 30;

    vr_status alloc(void** p, size_t bytes);
    // A block no queued work reads any more (its owner synchronized): kept for reuse.
    void release(void* p, size_t bytes);
    // A block queued work may still read: freed now (hipFree waits for the device).
    void release_now(void* p, size_t bytes);
    void trim();  // frees every cached block
    hipStream_t pick(void* s) const { return s ? (hipStream_t)s : stream; }
    void count(int64_t vr_stats::*field, int64_t v) {
        std::lock_guard<std::mutex> g(mu);
        stats.*field += v;
    }
};

// A device allocation owned by a handle; freed with the handle.
struct DevBuf {
    Device* dev = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : dev(o.dev), p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            dev = o.dev;
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    vr_status ensure(Device* d, size_t want) {  // grows (contents dropped) to at least `want` bytes
        if (want <= bytes && p) return VR_OK;
        if (p) dev->release_now(p, bytes);  // the old block may still be read by queued work
        p = nullptr;
        bytes = 0;
        dev = d;
        size_t b = want < 256 ? 256 : want;
        VR_TRY(d->alloc(&p, b));
        bytes = b;
        return VR_OK;
    }
    void reset() {  // the owner has synchronized: no queued work reads the block
        if (p) dev->release(p, bytes);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

// A page-locked host buffer for descriptors that a call writes and copies to the device: the copy
// is a DMA (no bounce through the runtime's own staging, as pageable memory takes), and the next
// call waits on the event recorded after the copy before it writes the buffer again.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipEvent_t copied = nullptr;
    bool pending = false;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() {
        if (pending) (void)hipEventSynchronize(copied);
        if (p) (void)hipHostFree(p);
        if (copied) (void)hipEventDestroy(copied);
    }
    // Waits for the last copy out of the buffer, then grows it (contents dropped) to `want` bytes.
    vr_status acquire(size_t want) {
        if (pending) {
            VR_TRY_HIP(hipEventSynchronize(copied));
            pending = false;
        }
        if (want <= bytes && p) return VR_OK;
        if (p) VR_TRY_HIP(hipHostFree(p));
        p = nullptr;
        bytes = 0;
        const size_t b = want + want / 4 + 4096;
        VR_TRY_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
        return VR_OK;
    }
    // Queues the copy of the first `n` bytes to `dst` on `s` and marks the buffer busy until it ends.
    vr_status copy_to(void* dst, size_t n, hipStream_t s) {
        if (!copied) VR_TRY_HIP(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
        VR_TRY_HIP(hipMemcpyAsync(dst, p, n, hipMemcpyHostToDevice, s));
        VR_TRY_HIP(hipEventRecord(copied, s));
        pending = true;
        return VR_OK;
    }
    template <class T>
    T* as() const { return (T*)p; }
};

// A host buffer to stage: its bytes and the device pointer to patch with its staged copy.
struct StagePiece {
    const void* src;
    size_t bytes;
    const void** dst;
};

// Pinned host staging: host columns are copied into a pinned buffer, then to the device on the
// side stream; the compute stream waits on the copy's event.  Two slots alternate so the copy of
// batch k+1 overlaps the kernels of batch k.  A large buffer (4 MiB or more) already in
// page-locked host memory is copied by DMA straight from the caller's buffer (no host copy).
struct Staging {
    Device* dev = nullptr;
    struct Slot {
        void* pinned = nullptr;
        size_t pinned_bytes = 0;
        DevBuf device;
        hipEvent_t copied = nullptr;    // H2D finished
        hipEvent_t consumed = nullptr;  // kernels reading the slot finished
        bool used = false;
    } slot[2];
    int next = 0;
    ~Staging();
    // Reserves `bytes` in the next slot: returns host (pinned) and device pointers.
    vr_status begin(Device* d, size_t bytes, uint8_t** host, uint8_t** device, int* which);
    // Marks the slot as read by the work queued on `compute` so far.
    vr_status done(int which, hipStream_t compute);
    // Stages the pieces through the next slot (256-byte aligned, in order), patches their device
    // pointers and makes `compute` wait for the copies; *which = -1 when there is nothing to stage.
    // With `defer` the copies are only queued: the caller calls ready() before the kernels that
    // read the slot, and keeps page-locked sources unchanged until the copy stream has drained.
    vr_status stage(Device* d, const std::vector<StagePiece>& pieces, hipStream_t compute, int* which,
                    bool defer = false);
    vr_status ready(int which, hipStream_t compute);
};

// True when `p` is page-locked host memory the DMA engine can read.
bool host_pinned(const void* p);

// Bytes of an Arrow buffer for n rows.
inline size_t bitmap_bytes(int64_t n) { return (size_t)((n + 7) / 8); }
inline size_t lane_bytes(int32_t type, int64_t n) {
    switch (type) {
        case VR_BOOL: return bitmap_bytes(n);
        case VR_INT32:
        case VR_UTF8: return (size_t)n * 4;
        case VR_INT64:
        case VR_FLOAT64: return (size_t)n * 8;
        case VR_DECIMAL128: return (size_t)n * 16;
    }
    return 0;
}
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace vr

struct vr_device {
    vr::Device d;
};

namespace vr {
void device_retain(vr_device* dev);
void device_release(vr_device* dev);

// Holds a reference on the device for a handle; declared before the handle's state so that the
// state's device buffers are freed first.
struct DeviceRef {
    vr_device* dev = nullptr;
    DeviceRef() = default;
    DeviceRef(const DeviceRef&) = delete;
    DeviceRef& operator=(const DeviceRef&) = delete;
    void hold(vr_device* d) {
        device_retain(d);
        dev = d;
    }
    ~DeviceRef() {
        if (dev) device_release(dev);
    }
};
}  // namespace vr
