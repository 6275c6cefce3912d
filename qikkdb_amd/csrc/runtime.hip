// runtime.hip - device binding, pooled allocation, per-stream scratch, copies and streams.
//
// Counterparts of the reference's GPU runtime substrate: Context (QueryEngine/Context.cu:20-127),
// CudaMemAllocator (QueryEngine/CudaMemAllocator.cu:109-322) and GPUMemory
// (QueryEngine/GPUCore/GPUMemory.cuh:89-189).  Differences by design: no single up-front
// reservation of all device memory (HBM is 288 GB; blocks are cached by size class instead),
// and every operation can run on a caller stream instead of the legacy default stream.
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace qk {

static thread_local char g_last_error[1024] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

// ---- device info ---------------------------------------------------------------------

static std::mutex g_info_mu;
static std::map<int, DeviceInfo> g_info;

int ensure_init() {
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_info_mu);
    if (g_info.count(dev)) return QK_SUCCESS;
    hipDeviceProp_t p;
    QK_HIP(hipGetDeviceProperties(&p, dev));
    DeviceInfo d;
    d.device = dev;
    d.cus = p.multiProcessorCount;
    d.lds_per_block = p.sharedMemPerBlock;
    d.lds_per_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor
                                                      : p.sharedMemPerBlock;
    d.max_threads = p.maxThreadsPerBlock;
    g_info[dev] = d;
    return QK_SUCCESS;
}

const DeviceInfo& device_info() {
    static thread_local DeviceInfo fallback;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fallback;
    {
        std::lock_guard<std::mutex> lk(g_info_mu);
        auto it = g_info.find(dev);
        if (it != g_info.end()) return it->second;
    }
    if (ensure_init() != QK_SUCCESS) return fallback;
    std::lock_guard<std::mutex> lk(g_info_mu);
    return g_info[dev];
}

// ---- pooled allocator -------------------------------------------------------------------

namespace {
struct Pool {
    std::mutex mu;
    std::multimap<size_t, void*> free_blocks;      // rounded size -> block
    std::unordered_map<void*, size_t> live;        // block -> rounded size
    size_t in_use = 0, cached = 0;
    size_t limit = 0;                              // bytes held from the driver at most (0: none)
};
std::mutex g_pools_mu;
std::map<int, Pool*> g_pools;

Pool& pool_for_current() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_pools_mu);
    Pool*& p = g_pools[dev];
    if (!p) p = new Pool();
    return *p;
}

size_t round_size(size_t b) {
    if (b == 0) b = 1;
    if (b < (1u << 20)) return (b + 255) & ~(size_t)255;
    const size_t g = (size_t)2 << 20;
    return (b + g - 1) / g * g;
}

int release_locked(Pool& p) {
    for (auto& kv : p.free_blocks) (void)hipFree(kv.second);
    p.free_blocks.clear();
    p.cached = 0;
    return QK_SUCCESS;
}
}  // namespace

int pool_alloc(void** out, size_t bytes) {
    if (!out) return fail(QK_ERROR, "qk_malloc: null output pointer");
    Pool& p = pool_for_current();
    size_t r = round_size(bytes);
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.free_blocks.find(r);
    if (it != p.free_blocks.end()) {
        *out = it->second;
        p.free_blocks.erase(it);
        p.cached -= r;
        p.live[*out] = r;
        p.in_use += r;
        return QK_SUCCESS;
    }
    if (p.limit) {
        // the cap covers what the pool holds from the driver: blocks in use and cached ones
        if (p.in_use + r > p.limit)
            return fail(QK_OUT_OF_MEMORY, "device allocation of %zu bytes exceeds the memory limit of %zu bytes "
                        "(%zu in use)", r, p.limit, p.in_use);
        if (p.in_use + p.cached + r > p.limit) release_locked(p);
    }
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, r);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_locked(p);  // evict the cache and retry, as GPUMemory::alloc does
        e = hipMalloc(&ptr, r);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            return fail(QK_OUT_OF_MEMORY, "device allocation of %zu bytes failed: %s", r,
                        hipGetErrorString(e));
        }
    }
    p.live[ptr] = r;
    p.in_use += r;
    *out = ptr;
    return QK_SUCCESS;
}

int pool_free(void* ptr) {
    if (!ptr) return QK_SUCCESS;
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    auto it = p.live.find(ptr);
    if (it == p.live.end()) return fail(QK_ERROR, "qk_free: %p was not allocated by qk_malloc", ptr);
    size_t r = it->second;
    p.live.erase(it);
    p.in_use -= r;
    p.free_blocks.emplace(r, ptr);
    p.cached += r;
    return QK_SUCCESS;
}

bool pool_can_allocate(size_t bytes) {
    Pool& p = pool_for_current();
    const size_t r = round_size(bytes);
    std::lock_guard<std::mutex> lk(p.mu);
    if (p.free_blocks.count(r)) return true;
    if (p.limit) return p.in_use + r <= p.limit;
    size_t free_b = 0, total = 0;
    if (hipMemGetInfo(&free_b, &total) != hipSuccess) return false;
    return free_b + p.cached >= r;
}

// ---- per-stream scratch -----------------------------------------------------------------

namespace {
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, Scratch*> g_scratch;
}  // namespace

int stream_scratch(hipStream_t s, size_t bytes, Scratch** out) {
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    Scratch*& sc = g_scratch[{dev, s}];
    if (!sc) sc = new Scratch();
    if (sc->bytes < bytes) {
        size_t nb = 1 << 16;
        while (nb < bytes) nb <<= 1;
        if (sc->ptr) {
            // the old buffer may still be read by work queued on this stream
            QK_HIP(hipStreamSynchronize(s));
            QK_HIP(hipFree(sc->ptr));
            sc->ptr = nullptr;
        }
        QK_HIP(hipMalloc(&sc->ptr, nb));
        QK_HIP(hipMemsetAsync(sc->ptr, 0, nb, s));
        sc->bytes = nb;
        sc->epoch = 0;
    }
    *out = sc;
    return QK_SUCCESS;
}

int pinned_host(size_t bytes, void** out) {
    static thread_local void* buf = nullptr;
    static thread_local size_t cap = 0;
    if (cap < bytes) {
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        size_t nb = 4096;
        while (nb < bytes) nb <<= 1;
        QK_HIP(hipHostMalloc(&buf, nb, hipHostMallocDefault));
        cap = nb;
    }
    *out = buf;
    return QK_SUCCESS;
}

}  // namespace qk

using namespace qk;

extern "C" {

int qk_init(int device) {
    int n = 0;
    QK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(QK_ERROR, "qk_init: device %d out of range (%d devices)", device, n);
    QK_HIP(hipSetDevice(device));
    QK_TRY(ensure_init());
    return prefix_noalloc_setup();
}

int qk_device_count(int* count) {
    if (!count) return fail(QK_ERROR, "qk_device_count: null output");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return fail(QK_RUNTIME_ERROR, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    return QK_SUCCESS;
}

const char* qk_last_error_message(void) { return g_last_error; }

int qk_malloc(void** ptr, size_t bytes) { return pool_alloc(ptr, bytes); }
int qk_free(void* ptr) { return pool_free(ptr); }

int qk_mem_stats(size_t* in_use, size_t* cached) {
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    if (in_use) *in_use = p.in_use;
    if (cached) *cached = p.cached;
    return QK_SUCCESS;
}

int qk_mem_release_cached(void) {
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    return release_locked(p);
}

int qk_mem_set_limit(size_t bytes) {
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    p.limit = bytes;
    if (p.limit && p.in_use + p.cached > p.limit) release_locked(p);
    return QK_SUCCESS;
}

int qk_mem_set_limit_percent(int32_t percent) {
    if (percent < 1 || percent > 100) return fail(QK_ERROR, "memory limit: percent must be in 1..100, got %d", percent);
    size_t free_b = 0, total = 0;
    QK_HIP(hipMemGetInfo(&free_b, &total));
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    // a share of what is free now plus what the pool already holds
    p.limit = (free_b + p.in_use + p.cached) / 100 * (size_t)percent;
    if (p.in_use + p.cached > p.limit) release_locked(p);
    return QK_SUCCESS;
}

int qk_mem_get_limit(size_t* bytes) {
    if (!bytes) return fail(QK_ERROR, "qk_mem_get_limit: null output");
    Pool& p = pool_for_current();
    std::lock_guard<std::mutex> lk(p.mu);
    *bytes = p.limit;
    return QK_SUCCESS;
}

int qk_mem_can_allocate(size_t bytes, int32_t* out) {
    if (!out) return fail(QK_ERROR, "qk_mem_can_allocate: null output");
    *out = pool_can_allocate(bytes) ? 1 : 0;
    return QK_SUCCESS;
}

int qk_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return QK_SUCCESS;
    QK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    QK_HIP(hipStreamSynchronize(as_stream(stream)));
    return QK_SUCCESS;
}

int qk_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return QK_SUCCESS;
    QK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    QK_HIP(hipStreamSynchronize(as_stream(stream)));
    return QK_SUCCESS;
}

int qk_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return QK_SUCCESS;
    QK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return QK_SUCCESS;
}

int qk_memset(void* dst, int byte_value, size_t bytes, void* stream) {
    if (bytes == 0) return QK_SUCCESS;
    QK_HIP(hipMemsetAsync(dst, byte_value, bytes, as_stream(stream)));
    return QK_SUCCESS;
}

int qk_stream_create(void** stream) {
    if (!stream) return fail(QK_ERROR, "qk_stream_create: null output");
    hipStream_t s;
    QK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return QK_SUCCESS;
}

int qk_stream_destroy(void* stream) {
    if (!stream) return QK_SUCCESS;
    QK_HIP(hipStreamDestroy(as_stream(stream)));
    return QK_SUCCESS;
}

int qk_stream_synchronize(void* stream) {
    QK_HIP(hipStreamSynchronize(as_stream(stream)));
    return QK_SUCCESS;
}

int qk_device_synchronize(void) {
    QK_HIP(hipDeviceSynchronize());
    return QK_SUCCESS;
}

}  // extern "C"
