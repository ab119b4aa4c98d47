// Device handle, errors, allocation accounting and pinned staging for libvrgpu.so.
#include <algorithm>
#include <cstring>
#include <thread>

#include "vr_internal.h"

namespace vr {

static thread_local std::string g_last_error;

vr_status fail(vr_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

vr_status hip_fail(hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(VR_ERR_OOM, std::string("device memory exhausted in ") + what);
    return fail(VR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

static size_t size_class(size_t bytes) {
    if (bytes <= (1u << 20)) {  // powers of two up to 1 MiB
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    size_t top = 1ull << (63 - __builtin_clzll(bytes));  // eighths of the power-of-two band above
    size_t step = top / 8;
    return (bytes + step - 1) / step * step;
}

vr_status Device::alloc(void** p, size_t bytes) {
    const size_t want = size_class(bytes);
    {
        std::lock_guard<std::mutex> g(mu);
        stats.device_allocations++;
        if (budget > 0 && allocated + (int64_t)want > budget)
            return fail(VR_ERR_OOM, "HBM budget of " + std::to_string(budget) + " bytes exceeded");
        auto it = cache.lower_bound(want);
        if (it != cache.end() && it->first <= want + want / 2) {
            *p = it->second;
            allocated += (int64_t)it->first;
            cached -= (int64_t)it->first;
            cache.erase(it);
            return VR_OK;
        }
        allocated += (int64_t)want;
    }
    hipError_t e = hipMalloc(p, want);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {  // give the cache back and retry
        (void)hipGetLastError();
        trim();
        e = hipMalloc(p, want);
    }
    std::lock_guard<std::mutex> g(mu);
    if (e != hipSuccess) {
        allocated -= (int64_t)want;
        (void)hipGetLastError();
        *p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    block_bytes[*p] = want;
    return VR_OK;
}

void Device::release(void* p, size_t) {
    if (!p) return;
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> g(mu);
        size_t b = block_bytes[p];
        allocated -= (int64_t)b;
        cache.emplace(b, p);
        cached += (int64_t)b;
        while (cached > cache_limit && !cache.empty()) {  // the largest blocks go first
            auto it = std::prev(cache.end());
            cached -= (int64_t)it->first;
            block_bytes.erase(it->second);
            drop.push_back(it->second);
            cache.erase(it);
        }
    }
    for (void* q : drop) (void)hipFree(q);
}

void Device::release_now(void* p, size_t) {
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> g(mu);
    allocated -= (int64_t)block_bytes[p];
    block_bytes.erase(p);
}

void Device::trim() {
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> g(mu);
        for (auto& kv : cache) {
            drop.push_back(kv.second);
            block_bytes.erase(kv.second);
        }
        cache.clear();
        cached = 0;
    }
    for (void* q : drop) (void)hipFree(q);
}

Staging::~Staging() {
    for (auto& s : slot) {
        if (s.pinned) (void)hipHostFree(s.pinned);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.consumed) (void)hipEventDestroy(s.consumed);
    }
}

vr_status Staging::begin(Device* d, size_t bytes, uint8_t** host, uint8_t** device, int* which) {
    dev = d;
    int w = next;
    next ^= 1;
    Slot& s = slot[w];
    if (!s.copied) {
        VR_TRY_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
        VR_TRY_HIP(hipEventCreateWithFlags(&s.consumed, hipEventDisableTiming));
    }
    if (s.used) VR_TRY_HIP(hipEventSynchronize(s.consumed));  // the kernels reading it are done
    if (s.pinned_bytes < bytes) {
        if (s.pinned) VR_TRY_HIP(hipHostFree(s.pinned));
        s.pinned = nullptr;
        size_t b = align_up(bytes + bytes / 4, 1 << 20);
        VR_TRY_HIP(hipHostMalloc(&s.pinned, b, hipHostMallocDefault));
        s.pinned_bytes = b;
    }
    VR_TRY(s.device.ensure(d, align_up(bytes + bytes / 4, 1 << 20)));
    *host = (uint8_t*)s.pinned;
    *device = s.device.as<uint8_t>();
    *which = w;
    return VR_OK;
}

bool host_pinned(const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is not known to the runtime
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

// Copies `n` bytes in up to `threads` host threads (pageable sources into the pinned buffer).
static void parallel_copies(const std::vector<std::pair<uint8_t*, const uint8_t*>>& dst_src,
                            const std::vector<size_t>& bytes, size_t total, int threads) {
    if (threads <= 1 || total < ((size_t)16 << 20)) {
        for (size_t q = 0; q < bytes.size(); q++) std::memcpy(dst_src[q].first, dst_src[q].second, bytes[q]);
        return;
    }
    // thread t copies bytes [t * total / threads, (t + 1) * total / threads) of the concatenation
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t] {
            const size_t lo = total * (size_t)t / (size_t)threads, hi = total * (size_t)(t + 1) / (size_t)threads;
            size_t at = 0;
            for (size_t q = 0; q < bytes.size() && at < hi; q++) {
                const size_t a = std::max(lo, at), b = std::min(hi, at + bytes[q]);
                if (a < b) std::memcpy(dst_src[q].first + (a - at), dst_src[q].second + (a - at), b - a);
                at += bytes[q];
            }
        });
    }
    for (auto& th : pool) th.join();
}

vr_status Staging::stage(Device* d, const std::vector<StagePiece>& pieces, hipStream_t compute, int* which, bool defer) {
    *which = -1;
    if (pieces.empty()) return VR_OK;
    size_t total = 0;
    for (auto& p : pieces) total = align_up(total, 256) + p.bytes;
    uint8_t *host, *device;
    VR_TRY(begin(d, total, &host, &device, which));
    Slot& s = slot[*which];
    // where each piece goes; a large page-locked source is copied by DMA as it is, every other
    // piece first into the pinned buffer (at its own offset, so consecutive pieces go in one copy:
    // a DMA per small piece costs more than the host copy)
    std::vector<size_t> offs(pieces.size());
    std::vector<char> pinned(pieces.size());
    std::vector<std::pair<uint8_t*, const uint8_t*>> dst_src;
    std::vector<size_t> sizes;
    size_t off = 0, pageable = 0;
    for (size_t q = 0; q < pieces.size(); q++) {
        off = align_up(off, 256);
        offs[q] = off;
        pinned[q] = pieces[q].bytes >= ((size_t)4 << 20) && host_pinned(pieces[q].src) ? 1 : 0;
        if (!pinned[q]) {
            dst_src.push_back({host + off, (const uint8_t*)pieces[q].src});
            sizes.push_back(pieces[q].bytes);
            pageable += pieces[q].bytes;
        }
        *pieces[q].dst = device + off;
        off += pieces[q].bytes;
    }
    parallel_copies(dst_src, sizes, pageable, (int)std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency())));
    bool direct = false;
    size_t run_lo = 0, run_hi = 0;  // a run of pinned-buffer pieces not copied yet
    auto flush = [&]() -> vr_status {
        if (run_hi > run_lo)
            VR_TRY_HIP(hipMemcpyAsync(device + run_lo, host + run_lo, run_hi - run_lo, hipMemcpyHostToDevice, dev->copy_stream));
        run_lo = run_hi = 0;
        return VR_OK;
    };
    for (size_t q = 0; q < pieces.size(); q++) {
        if (pinned[q]) {
            VR_TRY(flush());
            VR_TRY_HIP(hipMemcpyAsync(device + offs[q], pieces[q].src, pieces[q].bytes, hipMemcpyHostToDevice, dev->copy_stream));
            direct = true;
        } else {
            if (run_hi == run_lo) run_lo = offs[q];
            run_hi = offs[q] + pieces[q].bytes;
        }
    }
    VR_TRY(flush());
    dev->count(&vr_stats::h2d_bytes, (int64_t)total);
    VR_TRY_HIP(hipEventRecord(s.copied, dev->copy_stream));
    if (defer) return VR_OK;  // the caller waits (ready) and, for page-locked sources, synchronizes
    VR_TRY_HIP(hipStreamWaitEvent(compute, s.copied, 0));
    // the caller may reuse its buffers once the call returns: the copies out of them end first
    if (direct) VR_TRY_HIP(hipEventSynchronize(s.copied));
    return VR_OK;
}

vr_status Staging::ready(int which, hipStream_t compute) {
    VR_TRY_HIP(hipStreamWaitEvent(compute, slot[which].copied, 0));
    return VR_OK;
}

vr_status Staging::done(int which, hipStream_t compute) {
    Slot& s = slot[which];
    VR_TRY_HIP(hipEventRecord(s.consumed, compute));
    s.used = true;
    return VR_OK;
}

static void device_free(vr_device* dev) {
    (void)hipStreamSynchronize(dev->d.stream);
    (void)hipStreamSynchronize(dev->d.copy_stream);
    dev->d.trim();
    (void)hipStreamDestroy(dev->d.stream);
    (void)hipStreamDestroy(dev->d.copy_stream);
    delete dev;
}

void device_retain(vr_device* dev) { dev->d.refs.fetch_add(1); }

void device_release(vr_device* dev) {
    if (dev->d.refs.fetch_sub(1) == 1) device_free(dev);
}

}  // namespace vr

extern "C" {

const char* vr_last_error(void) { return vr::g_last_error.c_str(); }

vr_status vr_device_open(int32_t ordinal, int64_t hbm_budget_bytes, vr_device** out) {
    if (!out) return vr::fail(VR_ERR_INVALID_ARGUMENT, "null output handle");
    int count = 0;
    VR_TRY_HIP(hipGetDeviceCount(&count));
    if (ordinal < 0 || ordinal >= count)
        return vr::fail(VR_ERR_INVALID_ARGUMENT, "no device " + std::to_string(ordinal));
    VR_TRY_HIP(hipSetDevice(ordinal));
    auto* h = new vr_device();
    h->d.ordinal = ordinal;
    h->d.budget = hbm_budget_bytes;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) == hipSuccess) {
        h->d.num_cus = prop.multiProcessorCount;
        h->d.lds_per_block = (int64_t)prop.sharedMemPerBlock;
    }
    hipError_t e = hipStreamCreateWithFlags(&h->d.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->d.copy_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return vr::hip_fail(e, "hipStreamCreate");
    }
    *out = h;
    return VR_OK;
}

vr_status vr_device_close(vr_device* dev) {
    if (!dev) return VR_OK;
    vr::device_release(dev);  // freed now, or when its last handle is destroyed
    return VR_OK;
}

vr_status vr_device_memory(vr_device* dev, int64_t* allocated, int64_t* budget) {
    if (!dev) return vr::fail(VR_ERR_INVALID_ARGUMENT, "null device");
    std::lock_guard<std::mutex> g(dev->d.mu);
    if (allocated) *allocated = dev->d.allocated;
    if (budget) *budget = dev->d.budget;
    return VR_OK;
}

vr_status vr_device_sync(vr_device* dev, void* stream) {
    if (!dev) return vr::fail(VR_ERR_INVALID_ARGUMENT, "null device");
    if (stream) {
        VR_TRY_HIP(hipStreamSynchronize((hipStream_t)stream));
        return VR_OK;
    }
    VR_TRY_HIP(hipStreamSynchronize(dev->d.stream));
    VR_TRY_HIP(hipStreamSynchronize(dev->d.copy_stream));
    return VR_OK;
}

vr_status vr_stats_get(vr_device* dev, vr_stats* out) {
    if (!dev || !out) return vr::fail(VR_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> g(dev->d.mu);
    *out = dev->d.stats;
    return VR_OK;
}

}  // extern "C"
