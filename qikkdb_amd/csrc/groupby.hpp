// groupby.hpp - GROUP BY hash table layout shared by the table kernels and the host code.
#pragma once

#include "common.hpp"
#include "pred.hpp"

namespace qk {

constexpr int kMaxKeys = 8;
constexpr int kMaxAggs = 16;

// Empty marker of a key word in single-word tables.  A real key equal to it lives in a side
// slot instead, so every 64-bit key value is representable (the reference uses
// numeric_limits<K>::lowest() as the marker and cannot store that key, GPUGroupBy.cuh:23-30).
constexpr uint64_t kEmpty = 0x8000000000000000ull;
// Multi-word tables mark the last key word: 0 = empty, 1 = being written; a stored key always
// has bit 63 of its last word set.
constexpr uint64_t kBusy = 1ull;
constexpr uint64_t kPresent = 0x8000000000000000ull;

// aggregate state kept per group
enum : uint32_t { F_CNT = 1u, F_CNTALL = 2u, F_SUM = 4u, F_MIN = 8u, F_MAX = 16u };

// How a row's key columns become KW 64-bit words.
struct KeySpec {
    int32_t nkeys;
    int32_t kw;        // key words: 1, 2 or 4
    int32_t single;    // one key column, value widened to 64 bits, NULL key -> side slot
    int32_t nullable;  // multi-word layout carries one null flag per column
    int32_t type[kMaxKeys];
    int32_t word[kMaxKeys], shift[kMaxKeys];    // value field position
    int32_t nword[kMaxKeys], nshift[kMaxKeys];  // null flag position
};

struct KeyCols {
    const void* col[kMaxKeys];
    const uint64_t* nm[kMaxKeys];
};

// Record layout in 64-bit words; -1 marks an absent field.
struct Rec {
    int32_t words, kw, cnt, cntall, sum, mn, mx;
};

// Device view of one table.  Slots [0, cap) are hashed; slot cap holds the NULL key of a
// single-key table, slot cap + 1 the key equal to kEmpty.
struct GView {
    uint64_t* slots;
    uint64_t cap;
    Rec rec;
    unsigned long long* nkeys;  // distinct non-NULL keys inserted
    int32_t* err;
    uint64_t logical_cap;
    int32_t real_sum;  // SUM state is FP64 (else wrapping int64)
};

struct TableCounters {
    unsigned long long nkeys;
    int32_t err;
    int32_t pad;
};

__device__ __host__ __forceinline__ uint64_t* rec_ptr(const GView& g, uint64_t s) {
    return g.slots + s * (uint64_t)g.rec.words;
}

}  // namespace qk
