// GPUMemoryCache.h - the block cache of the C++ drop-in layer, with the call shapes of
// QueryEngine/GPUMemoryCache.h:10-141 of the reference: GetColumn<T> returns (pointer, element
// count, cache hit); Evict, ClearCachedBlock, ContainsColumn, SetCurrentBlockIndex and the static
// SetLockList.  The blocks live in the library's pool (qk_cache_* in qikkdb_amd.h), under the
// pool's cap (qk_mem_set_limit, the counterpart of GPUMemoryUsagePercent).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "GPUError.h"

class GPUMemoryCache {
    qk_cache cache_ = nullptr;

public:
    /// Create a cache of maximumSize bytes on device deviceID (GPUMemoryCache.cpp:13-17).
    GPUMemoryCache(int32_t deviceID, size_t maximumSize) {
        qikkdb::check(qk_init(deviceID));
        qikkdb::check(qk_cache_create(&cache_, maximumSize));
    }
    ~GPUMemoryCache() { qk_cache_destroy(cache_); }
    GPUMemoryCache(const GPUMemoryCache&) = delete;
    GPUMemoryCache& operator=(const GPUMemoryCache&) = delete;

    /// Cached block, or a fresh one (contents undefined) evicting least recently used blocks
    /// first; throws std::length_error when the block is larger than the cache or nothing is
    /// left to evict (GPUMemoryCache.h:98-140).
    template <typename T>
    std::tuple<T*, size_t, bool> GetColumn(const std::string& databaseName, const std::string& tableAndColumnName,
                                           int32_t blockIndex, size_t size, int64_t loadSize, int64_t loadOffset) {
        void* p = nullptr;
        size_t bytes = 0;
        int32_t hit = 0;
        const int st = qk_cache_get_column(cache_, databaseName.c_str(), tableAndColumnName.c_str(), blockIndex,
                                           sizeof(T) * size, loadSize, loadOffset, &p, &bytes, &hit);
        if (st == QK_ERROR || st == QK_OUT_OF_MEMORY) throw std::length_error(qk_last_error_message());
        qikkdb::check(st);
        return std::tuple<T*, size_t, bool>{static_cast<T*>(p), hit ? bytes / sizeof(T) : size, hit != 0};
    }

    /// Evict the least recently used block not on the lock list; false when there is none.
    bool Evict() {
        int32_t evicted = 0;
        qikkdb::check(qk_cache_evict(cache_, &evicted));
        return evicted != 0;
    }

    void ClearCachedBlock(const std::string& databaseName, const std::string& tableAndColumnName, int32_t blockIndex) {
        qikkdb::check(qk_cache_clear_block(cache_, databaseName.c_str(), tableAndColumnName.c_str(), blockIndex));
    }

    bool ContainsColumn(const std::string& databaseName, const std::string& tableAndColumnName, int32_t blockIndex,
                        int64_t loadSize, int64_t loadOffset) {
        int32_t out = 0;
        qikkdb::check(qk_cache_contains(cache_, databaseName.c_str(), tableAndColumnName.c_str(), blockIndex,
                                        loadSize, loadOffset, &out));
        return out != 0;
    }

    void SetCurrentBlockIndex(size_t blockIndex) {
        qikkdb::check(qk_cache_set_current_block(cache_, static_cast<int64_t>(blockIndex)));
    }

    /// Columns ("database.table.column") whose blocks of the current block index are never
    /// evicted; one list for every cache, as in the reference.
    static void SetLockList(const std::vector<std::string>& lockList) {
        std::vector<const char*> v;
        for (const std::string& s : lockList) v.push_back(s.c_str());
        qikkdb::check(qk_cache_set_lock_list(static_cast<int32_t>(v.size()), v.empty() ? nullptr : v.data()));
    }
};
