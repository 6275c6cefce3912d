// Host test of the block cache's bookkeeping (qikkdb_amd/csrc/cache.hpp) over a host
// allocator with a byte cap: hits and misses, least-recently-used eviction, the lock list,
// removal of one block, refusal of a block larger than the cache, eviction forced by the
// allocator's cap.  Prints "cache ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstdlib>
#include <set>

#include "../../qikkdb_amd/csrc/cache.hpp"

namespace {

struct HostAlloc {
    size_t cap = 0, used = 0;
    std::set<void*> live;
    int alloc(void** p, size_t bytes) {
        if (cap && used + bytes > cap) return 65;
        *p = std::malloc(bytes ? bytes : 1);
        live.insert(*p);
        used += bytes;
        sizes[*p] = bytes;
        return 0;
    }
    void release(void* p) {
        used -= sizes[p];
        sizes.erase(p);
        live.erase(p);
        std::free(p);
    }
    bool can_allocate(size_t bytes) const { return !cap || used + bytes <= cap; }
    void before_evict() { rounds++; }
    int rounds = 0;
    std::map<void*, size_t> sizes;
};

int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAILED line %d: %s\n", __LINE__, #c);          \
            failures++;                                                 \
        }                                                               \
    } while (0)

using Cache = qk::BlockCache<HostAlloc>;

bool get(Cache& c, const char* col, int block, size_t bytes, void** p = nullptr, int* rc = nullptr) {
    void* ptr = nullptr;
    size_t got = 0;
    bool hit = false;
    int r = c.get("db", col, block, bytes, -1, 0, &ptr, &got, &hit);
    if (rc) *rc = r;
    if (p) *p = ptr;
    return hit;
}

}  // namespace

int main() {
    {  // hits, misses, LRU order
        HostAlloc al;
        Cache c(al, 300);
        void *a = nullptr, *a2 = nullptr;
        CHECK(!get(c, "t.a", 0, 100, &a));
        CHECK(get(c, "t.a", 0, 100, &a2) && a2 == a);
        CHECK(!get(c, "t.b", 0, 100));
        CHECK(!get(c, "t.c", 0, 100));
        CHECK(get(c, "t.a", 0, 100));  // a is now the most recent: b is the oldest
        CHECK(!get(c, "t.d", 0, 100));  // evicts b
        CHECK(!c.contains("db", "t.b", 0, -1, 0));
        CHECK(c.contains("db", "t.a", 0, -1, 0) && c.contains("db", "t.c", 0, -1, 0) && c.contains("db", "t.d", 0, -1, 0));
        const auto s = c.stats();
        CHECK(s.used == 300 && s.entries == 3 && s.hits == 2 && s.misses == 4 && s.evictions == 1);
        CHECK(al.used == 300 && al.live.size() == 3 && al.rounds == 1);
        // load size and offset are part of the name
        void* p = nullptr;
        size_t got = 0;
        bool hit = true;
        CHECK(c.get("db", "t.a", 0, 50, 1000, 64, &p, &got, &hit) == Cache::kOk && !hit);
    }
    {  // the lock list protects blocks of the current block only
        HostAlloc al;
        Cache c(al, 200);
        c.set_lock_list({"db.t.a"});
        c.set_current_block(1);
        get(c, "t.a", 1, 100);
        get(c, "t.b", 1, 100);
        CHECK(!get(c, "t.c", 1, 100));  // a is locked and oldest: b goes
        CHECK(c.contains("db", "t.a", 1, -1, 0) && !c.contains("db", "t.b", 1, -1, 0));
        c.set_current_block(2);
        CHECK(!get(c, "t.d", 2, 100));  // a (block 1) is no longer locked and is the oldest
        CHECK(!c.contains("db", "t.a", 1, -1, 0));
        // everything locked: no room
        c.set_lock_list({"db.t.c", "db.t.d"});
        c.set_current_block(2);
        get(c, "t.c", 2, 100);
        get(c, "t.d", 2, 100);
        int rc = -1;
        get(c, "t.e", 2, 100, nullptr, &rc);
        CHECK(rc == Cache::kNoRoom);
    }
    {  // removal of one block, whatever its load size; block 12 survives a clear of block 1
        HostAlloc al;
        Cache c(al, 1000);
        void* p = nullptr;
        size_t got = 0;
        bool hit = false;
        c.get("db", "t.a", 1, 10, -1, 0, &p, &got, &hit);
        c.get("db", "t.a", 1, 10, 5, 5, &p, &got, &hit);
        c.get("db", "t.a", 12, 10, -1, 0, &p, &got, &hit);
        c.get("db", "t.ab", 1, 10, -1, 0, &p, &got, &hit);
        c.clear_block("db", "t.a", 1);
        CHECK(!c.contains("db", "t.a", 1, -1, 0) && !c.contains("db", "t.a", 1, 5, 5));
        CHECK(c.contains("db", "t.a", 12, -1, 0) && c.contains("db", "t.ab", 1, -1, 0));
        CHECK(c.stats().entries == 2 && c.stats().used == 20 && al.used == 20);
    }
    {  // too large; the allocator's cap forces eviction below the cache's own budget
        HostAlloc al;
        al.cap = 250;
        Cache c(al, 1000);
        int rc = -1;
        get(c, "t.big", 0, 1001, nullptr, &rc);
        CHECK(rc == Cache::kTooLarge);
        get(c, "t.a", 0, 100);
        get(c, "t.b", 0, 100);
        get(c, "t.c", 0, 100, nullptr, &rc);  // 300 > 250: a goes
        CHECK(rc == Cache::kOk && !c.contains("db", "t.a", 0, -1, 0) && al.used == 200);
        // memory the cache does not own counts against the cap too
        al.cap = 400;
        void* other = nullptr;
        CHECK(al.alloc(&other, 150) == 0);
        get(c, "t.d", 0, 100, nullptr, &rc);  // 450 > 400: b goes
        CHECK(rc == Cache::kOk && c.stats().entries == 2 && al.used == 350);
        get(c, "t.e", 0, 300, nullptr, &rc);  // 150 held outside: never fits, the cache empties
        CHECK(rc == Cache::kNoRoom && c.stats().entries == 0 && al.used == 150);
        al.release(other);
    }
    {  // destruction frees every block
        HostAlloc al;
        {
            Cache c(al, 1000);
            get(c, "t.a", 0, 10);
            get(c, "t.b", 0, 10);
        }
        CHECK(al.used == 0 && al.live.empty());
    }
    if (failures) {
        std::printf("%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("cache ok\n");
    return 0;
}
