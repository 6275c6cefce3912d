// cache.hpp - least-recently-used cache of device column blocks.
//
// Counterpart of GPUMemoryCache (QueryEngine/GPUMemoryCache.h:10-141, GPUMemoryCache.cpp:31-117):
// a block is named by database, table.column, block index, load size and load offset; a hit
// moves it to the most recently used end, a miss allocates it, evicting least recently used
// blocks until the cache's byte budget and the allocator both have room.  Blocks of the current
// block index whose column is on the lock list are never evicted.
//
// The bookkeeping is written over an allocator interface so that it runs on the host as well
// (the library instantiates it over its device pool):
//     int  alloc(void** p, size_t bytes);     // QK_SUCCESS or an error code
//     void release(void* p);
//     bool can_allocate(size_t bytes) const;  // would alloc succeed under the allocator's cap
//     void before_evict();                    // called once before a call's first eviction
//
// Differences from the reference, by design: ClearCachedBlock removes exactly the named block
// (the reference's prefix match on "db.table.column_<block>" also removes block 12 when asked
// for block 1), and a column is locked when its name equals a lock-list entry and the entry's
// block is the current block (the reference compares the last digit position of the key).
#pragma once

#include <cstddef>
#include <cstdint>
#include <list>
#include <map>
#include <string>
#include <vector>

namespace qk {

template <class Alloc>
class BlockCache {
public:
    struct Stats {
        size_t used = 0;
        int64_t entries = 0, hits = 0, misses = 0, evictions = 0;
    };

    BlockCache(Alloc& al, size_t max_bytes) : al_(al), max_(max_bytes) {}
    ~BlockCache() { clear(); }
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;

    static std::string column_of(const std::string& db, const std::string& table_column) {
        return db + "." + table_column;
    }
    static std::string key_of(const std::string& column, int32_t block, int64_t load_size, int64_t load_offset) {
        return column + "_" + std::to_string(block) + "_" + std::to_string(load_size) + "_" +
               std::to_string(load_offset);
    }

    // Outcome of get(): 0 = ok, 1 = the block is larger than the whole cache, 2 = nothing left
    // to evict and still no room, 3 = the allocator failed (its code in *alloc_status).
    enum { kOk = 0, kTooLarge = 1, kNoRoom = 2, kAllocFailed = 3 };

    int get(const std::string& db, const std::string& table_column, int32_t block, size_t bytes,
            int64_t load_size, int64_t load_offset, void** ptr, size_t* out_bytes, bool* hit,
            int* alloc_status = nullptr) {
        const std::string column = column_of(db, table_column);
        const std::string key = key_of(column, block, load_size, load_offset);
        auto it = map_.find(key);
        if (it != map_.end()) {
            touch(it->second);
            *ptr = it->second.ptr;
            *out_bytes = it->second.bytes;
            *hit = true;
            st_.hits++;
            return kOk;
        }
        st_.misses++;
        *hit = false;
        if (bytes > max_) return kTooLarge;
        while (used_ + bytes > max_ || !al_.can_allocate(bytes))
            if (!evict()) {
                synced_ = false;
                return kNoRoom;
            }
        void* p = nullptr;
        int rc = al_.alloc(&p, bytes);
        // the allocator may hold memory the cap check did not see: evict and retry until the
        // cache is empty
        while (rc != 0 && evict()) rc = al_.alloc(&p, bytes);
        synced_ = false;
        if (rc != 0) {
            if (alloc_status) *alloc_status = rc;
            return kAllocFailed;
        }
        lru_.push_back(key);
        Entry e;
        e.column = column;
        e.block = block;
        e.ptr = p;
        e.bytes = bytes;
        e.pos = std::prev(lru_.end());
        map_.emplace(key, e);
        used_ += bytes;
        *ptr = p;
        *out_bytes = bytes;
        return kOk;
    }

    bool contains(const std::string& db, const std::string& table_column, int32_t block, int64_t load_size,
                  int64_t load_offset) const {
        return map_.count(key_of(column_of(db, table_column), block, load_size, load_offset)) != 0;
    }

    // Evicts the least recently used block that is not locked; false when there is none.
    bool evict() {
        for (auto it = lru_.begin(); it != lru_.end(); ++it) {
            auto m = map_.find(*it);
            if (locked(m->second)) continue;
            if (!synced_) {
                al_.before_evict();
                synced_ = true;
            }
            drop(m);
            st_.evictions++;
            return true;
        }
        return false;
    }

    // Whether any block of (db, table_column, block) is cached, whatever its load size and offset.
    bool contains_block(const std::string& db, const std::string& table_column, int32_t block) const {
        const std::string column = column_of(db, table_column);
        for (const auto& kv : map_)
            if (kv.second.column == column && kv.second.block == block) return true;
        return false;
    }

    // Removes every block of (db, table_column, block), whatever its load size and offset.
    void clear_block(const std::string& db, const std::string& table_column, int32_t block) {
        const std::string column = column_of(db, table_column);
        for (auto it = map_.begin(); it != map_.end();) {
            auto next = std::next(it);
            if (it->second.column == column && it->second.block == block) drop(it);
            it = next;
        }
    }

    void clear() {
        while (!map_.empty()) drop(map_.begin());
    }

    // Ends an eviction round: the next eviction calls before_evict again.
    void end_round() { synced_ = false; }
    void set_lock_list(const std::vector<std::string>& columns) { locks_ = columns; }
    void set_current_block(int64_t block) { current_ = block; }

    Stats stats() const {
        Stats s = st_;
        s.used = used_;
        s.entries = (int64_t)map_.size();
        return s;
    }
    size_t max_bytes() const { return max_; }

private:
    struct Entry {
        std::string column;
        int32_t block = 0;
        void* ptr = nullptr;
        size_t bytes = 0;
        std::list<std::string>::iterator pos;
    };

    void touch(Entry& e) { lru_.splice(lru_.end(), lru_, e.pos); }

    bool locked(const Entry& e) const {
        if (e.block != current_) return false;
        for (const std::string& c : locks_)
            if (c == e.column) return true;
        return false;
    }

    void drop(typename std::map<std::string, Entry>::iterator it) {
        al_.release(it->second.ptr);
        used_ -= it->second.bytes;
        lru_.erase(it->second.pos);
        map_.erase(it);
    }

    Alloc& al_;
    size_t max_;
    size_t used_ = 0;
    int64_t current_ = 0;
    bool synced_ = false;
    std::vector<std::string> locks_;
    std::map<std::string, Entry> map_;
    std::list<std::string> lru_;  // front: least recently used
    Stats st_;
};

}  // namespace qk
