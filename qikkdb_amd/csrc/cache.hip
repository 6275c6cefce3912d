// cache.hip - the block cache of device column blocks (qk_cache_* in include/qikkdb_amd.h), the
// counterpart of GPUMemoryCache (QueryEngine/GPUMemoryCache.h:10-141, GPUMemoryCache.cpp:13-117)
// over the library's pool (whose cap, qk_mem_set_limit, plays the part of the reference
// allocator's GPUMemoryUsagePercent reservation, CudaMemAllocator.cu:109-152).  Host code only.
#include <mutex>
#include <string>
#include <vector>

#include "cache.hpp"
#include "common.hpp"

namespace {

// The pool as the cache's allocator.  Blocks go back to the pool at once, so before the first
// eviction of a call the device is synchronised: no queued kernel still reads an evicted block
// when the pool hands its memory out again.
struct PoolAlloc {
    int alloc(void** p, size_t bytes) { return qk::pool_alloc(p, bytes); }
    void release(void* p) { qk::pool_free(p); }
    bool can_allocate(size_t bytes) const { return qk::pool_can_allocate(bytes); }
    void before_evict() { (void)hipDeviceSynchronize(); }
};

// The lock list is static in the reference (one list for every device's cache).
std::mutex g_lock_mu;
std::vector<std::string> g_locks;
uint64_t g_lock_version = 1;

}  // namespace

struct qk_cache_s {
    int device = 0;
    PoolAlloc al;
    qk::BlockCache<PoolAlloc> c;
    std::mutex mu;
    uint64_t lock_version = 0;
    explicit qk_cache_s(size_t max_bytes) : c(al, max_bytes) {}
};

namespace {

// Binds the cache's device and brings its copy of the lock list up to date (cache lock held).
int enter(qk_cache_s* c) {
    QK_HIP(hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(g_lock_mu);
    if (c->lock_version != g_lock_version) {
        c->c.set_lock_list(g_locks);
        c->lock_version = g_lock_version;
    }
    return QK_SUCCESS;
}

}  // namespace

extern "C" {

int qk_cache_create(qk_cache* out, size_t max_bytes) {
    if (!out) return qk::fail(QK_ERROR, "qk_cache_create: null output");
    QK_TRY(qk::ensure_init());
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    qk_cache_s* c = new qk_cache_s(max_bytes);
    c->device = dev;
    *out = c;
    return QK_SUCCESS;
}

int qk_cache_destroy(qk_cache c) {
    if (!c) return QK_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        QK_HIP(hipSetDevice(c->device));
        QK_HIP(hipDeviceSynchronize());
        c->c.clear();
    }
    delete c;
    return QK_SUCCESS;
}

int qk_cache_get_column(qk_cache c, const char* database, const char* table_column, int32_t block_index,
                        size_t bytes, int64_t load_size, int64_t load_offset, void** ptr, size_t* out_bytes,
                        int32_t* hit) {
    if (!c || !database || !table_column || !ptr || !out_bytes || !hit)
        return qk::fail(QK_ERROR, "qk_cache_get_column: null argument");
    if (bytes == 0) return qk::fail(QK_ERROR, "qk_cache_get_column: zero-byte block");
    std::lock_guard<std::mutex> lk(c->mu);
    QK_TRY(enter(c));
    bool h = false;
    int ast = QK_SUCCESS;
    const int rc = c->c.get(database, table_column, block_index, bytes, load_size, load_offset, ptr, out_bytes, &h, &ast);
    *hit = h ? 1 : 0;
    switch (rc) {
    case qk::BlockCache<PoolAlloc>::kOk: return QK_SUCCESS;
    case qk::BlockCache<PoolAlloc>::kTooLarge:
        return qk::fail(QK_ERROR, "Tried to cache block larger than maximum cache size (%zu > %zu bytes)", bytes,
                        c->c.max_bytes());
    case qk::BlockCache<PoolAlloc>::kNoRoom:
        return qk::fail(QK_OUT_OF_MEMORY, "Not enough space left in cache for %zu bytes", bytes);
    default: return ast;  // the pool's own message stands
    }
}

int qk_cache_contains(qk_cache c, const char* database, const char* table_column, int32_t block_index,
                      int64_t load_size, int64_t load_offset, int32_t* out) {
    if (!c || !database || !table_column || !out) return qk::fail(QK_ERROR, "qk_cache_contains: null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    *out = c->c.contains(database, table_column, block_index, load_size, load_offset) ? 1 : 0;
    return QK_SUCCESS;
}

int qk_cache_evict(qk_cache c, int32_t* evicted) {
    if (!c || !evicted) return qk::fail(QK_ERROR, "qk_cache_evict: null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    QK_TRY(enter(c));
    *evicted = c->c.evict() ? 1 : 0;
    c->c.end_round();
    return QK_SUCCESS;
}

int qk_cache_clear_block(qk_cache c, const char* database, const char* table_column, int32_t block_index) {
    if (!c || !database || !table_column) return qk::fail(QK_ERROR, "qk_cache_clear_block: null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    QK_TRY(enter(c));
    if (c->c.contains_block(database, table_column, block_index)) QK_HIP(hipDeviceSynchronize());
    c->c.clear_block(database, table_column, block_index);
    return QK_SUCCESS;
}

int qk_cache_set_current_block(qk_cache c, int64_t block_index) {
    if (!c) return qk::fail(QK_ERROR, "qk_cache_set_current_block: null cache");
    std::lock_guard<std::mutex> lk(c->mu);
    c->c.set_current_block(block_index);
    return QK_SUCCESS;
}

int qk_cache_set_lock_list(int32_t n, const char* const* columns) {
    if (n < 0 || (n > 0 && !columns)) return qk::fail(QK_ERROR, "qk_cache_set_lock_list: bad argument");
    std::vector<std::string> v;
    for (int32_t i = 0; i < n; i++) {
        if (!columns[i]) return qk::fail(QK_ERROR, "qk_cache_set_lock_list: null column %d", i);
        v.emplace_back(columns[i]);
    }
    std::lock_guard<std::mutex> lk(g_lock_mu);
    g_locks = std::move(v);
    g_lock_version++;
    return QK_SUCCESS;
}

int qk_cache_stats(qk_cache c, size_t* used, int64_t* entries, int64_t* hits, int64_t* misses, int64_t* evictions) {
    if (!c) return qk::fail(QK_ERROR, "qk_cache_stats: null cache");
    std::lock_guard<std::mutex> lk(c->mu);
    const auto s = c->c.stats();
    if (used) *used = s.used;
    if (entries) *entries = s.entries;
    if (hits) *hits = s.hits;
    if (misses) *misses = s.misses;
    if (evictions) *evictions = s.evictions;
    return QK_SUCCESS;
}

}  // extern "C"
