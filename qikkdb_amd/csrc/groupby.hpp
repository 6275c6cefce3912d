// groupby.hpp - GROUP BY hash table layout and device helpers shared by the table kernels
// (groupby.hip) and the two-phase partitioned kernels (twophase.hip).
#pragma once

#include <type_traits>

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

// Device view of one table, stored as arrays (structure of arrays): key words of slot s at
// keys[s*kw .. s*kw+kw), one array per aggregate state (nullptr when absent).  Slots
// [0, cap) are hashed; slot cap holds the NULL key of a single-word table, slot cap + 1 the
// key equal to kEmpty.  The hashed slots form `parts` sub-tables of cap/parts slots: a key
// lives in sub-table (hash >> part_shift) and probes stay inside it, so one workgroup can own
// a sub-table.
struct GView {
    uint64_t* keys;
    unsigned long long* cnt;
    unsigned long long* cntall;
    unsigned long long* sum;  // FP64 bits or wrapping int64
    unsigned long long* mn;
    unsigned long long* mx;
    uint64_t cap;
    uint64_t sub_mask;        // cap/parts - 1
    int32_t part_shift;       // 64 - log2(parts) (64 when parts == 1)
    int32_t sub_bits;         // log2(cap/parts)
    int32_t kw;
    int32_t real_sum;
    unsigned long long* nkeys;  // distinct non-NULL keys inserted
    int32_t* err;
    uint64_t logical_cap;
    int32_t* lost;              // set when a partial found no slot (its rows were dropped)
};

struct TableCounters {
    unsigned long long nkeys;
    int32_t err;
    int32_t lost;
};

// Partial aggregate of one row, one LDS slot or one merge record.
struct Part {
    uint64_t cnt, cntall, sum, mn, mx;  // sum holds FP64 bits or wrapping int64
};

__device__ __forceinline__ uint32_t part_of(const GView& g, uint64_t h) {
    return g.part_shift >= 64 ? 0u : (uint32_t)(h >> g.part_shift);
}
__device__ __forceinline__ uint64_t home_of(const GView& g, uint64_t h) {
    return ((uint64_t)part_of(g, h) << g.sub_bits) | (h & g.sub_mask);
}
__device__ __forceinline__ uint64_t next_in_sub(const GView& g, uint64_t s) {
    return (s & ~g.sub_mask) | ((s + 1) & g.sub_mask);
}

// ---- key normalisation -----------------------------------------------------------------

// Bits of a FLOAT/DOUBLE key with -0.0 folded into +0.0 (so equal values group together) and
// every NaN (either sign, any payload) folded into the default quiet NaN, so that NaN keys form
// one group.  The reference compares keys with `!=` (GPUGroupBy.cuh:105), under which a NaN
// key matches no slot and each NaN row opens a group of its own.
__device__ __forceinline__ uint64_t fkey32(float f) {
    const uint32_t b = __float_as_uint(f);
    const uint32_t m = b & 0x7fffffffu;
    return m > 0x7f800000u ? 0x7fc00000ull : (m ? (uint64_t)b : 0ull);
}
__device__ __forceinline__ uint64_t fkey64(double d) {
    const uint64_t b = (uint64_t)__double_as_longlong(d);
    const uint64_t m = b & 0x7fffffffffffffffull;
    return m > 0x7ff0000000000000ull ? 0x7ff8000000000000ull : (m ? b : 0ull);
}

// Single key column widened to one 64-bit word.
__device__ __forceinline__ uint64_t key_single(int type, const void* col, int64_t r) {
    switch (type) {
    case QK_INT: return (uint64_t)(int64_t)((const int32_t*)col)[r];
    case QK_LONG: return (uint64_t)((const int64_t*)col)[r];
    case QK_FLOAT: return fkey32(((const float*)col)[r]);
    case QK_DOUBLE: return fkey64(((const double*)col)[r]);
    default: return (uint64_t)(int64_t)((const int8_t*)col)[r];
    }
}

// Field bits of a composite key column: 32-bit fields for INT/FLOAT/INT8, 64-bit otherwise.
__device__ __forceinline__ uint64_t key_field(int type, const void* col, int64_t r) {
    switch (type) {
    case QK_INT: return (uint64_t)(uint32_t)((const int32_t*)col)[r];
    case QK_LONG: return (uint64_t)((const int64_t*)col)[r];
    case QK_FLOAT: return fkey32(((const float*)col)[r]);
    case QK_DOUBLE: return fkey64(((const double*)col)[r]);
    default: return (uint64_t)(uint32_t)(int32_t)((const int8_t*)col)[r];
    }
}

template <int KW>
__device__ __forceinline__ void load_key(const KeySpec& ks, const KeyCols& kc, int64_t r,
                                         uint64_t (&k)[KW], bool& knull) {
    knull = false;
    if (KW == 1 && ks.single) {
        knull = kc.nm[0] && null_bit(kc.nm[0], r);
        k[0] = knull ? 0ull : key_single(ks.type[0], kc.col[0], r);
        return;
    }
#pragma unroll
    for (int w = 0; w < KW; w++) k[w] = 0;
    for (int c = 0; c < ks.nkeys; c++) {
        const bool isnull = kc.nm[c] && null_bit(kc.nm[c], r);
        if (isnull) {
            if (KW > 1) {
#pragma unroll
                for (int w = 0; w < KW; w++)
                    if (w == ks.nword[c]) k[w] |= 1ull << ks.nshift[c];
            }
        } else {
            const uint64_t f = key_field(ks.type[c], kc.col[c], r) << ks.shift[c];
#pragma unroll
            for (int w = 0; w < KW; w++)
                if (w == ks.word[c]) k[w] |= f;
        }
    }
    if (KW > 1) k[KW - 1] |= kPresent;
}

template <int KW>
__device__ __forceinline__ uint64_t key_hash(const uint64_t (&k)[KW]) {
    uint64_t h = hash64(k[0]);
#pragma unroll
    for (int w = 1; w < KW; w++) h = hash64(h ^ (k[w] + 0x9e3779b97f4a7c15ull * (uint64_t)w));
    return h;
}

// ---- values --------------------------------------------------------------------------

struct NoValue {};

template <typename VT> struct VTraits {
    static constexpr bool has = true;
    static constexpr bool real = std::is_floating_point<VT>::value;
};
template <> struct VTraits<NoValue> {
    static constexpr bool has = false;
    static constexpr bool real = false;
};

// Partial of one row (value already known non-NULL unless vnull).
template <typename VT>
__device__ __forceinline__ Part row_part(const VT& v, bool vnull) {
    Part p;
    p.cntall = 1;
    p.cnt = vnull ? 0 : 1;
    p.sum = 0;
    p.mn = ~0ull;
    p.mx = 0;
    if constexpr (VTraits<VT>::has) {
        if (!vnull) {
            if constexpr (VTraits<VT>::real) {
                const double d = (double)v;
                p.sum = (uint64_t)__double_as_longlong(d);
                p.mn = p.mx = enc_f64(d);
            } else {
                const int64_t iv = (int64_t)v;
                p.sum = (uint64_t)iv;
                p.mn = p.mx = enc_i64(iv);
            }
        }
    }
    return p;
}

// ---- global table: probing with atomics (any number of writers) --------------------------

__device__ __forceinline__ void g_set_error(const GView& g, int32_t code) {
    __hip_atomic_store(to_g(g.err), code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A partial whose key found no free slot in its sub-table: the table reports itself full and
// records that rows were dropped.  A sub-table never loses a key once it holds it, so the
// dropped partials are exactly those whose key is absent from the table afterwards (every
// partial of a key the table holds was aggregated); a growing table re-aggregates those rows.
__device__ __forceinline__ void g_set_lost(const GView& g) {
    __hip_atomic_store(to_g(g.lost), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    g_set_error(g, QK_HASH_TABLE_FULL);
}

// Side slot claim (0 -> 1); returns true when this call claimed it.
__device__ __forceinline__ bool g_claim_side(const GView& g, uint64_t s) {
    uint64_t* kp = g.keys + s * (uint64_t)g.kw;
    if (g_load_volatile(kp) != 0ull) return false;
    return g_atomic_cas((unsigned long long*)kp, 0ull, 1ull) == 0ull;
}

// Slot of a single-word key (inserting it when absent); -1 when its sub-table is full.
// *is_new tells the caller to count the key.
__device__ __forceinline__ int64_t g_slot1(const GView& g, uint64_t key, bool knull, uint64_t h,
                                           bool* is_new) {
    *is_new = false;
    if (knull) {
        g_claim_side(g, g.cap);
        return (int64_t)g.cap;
    }
    if (key == kEmpty) {
        *is_new = g_claim_side(g, g.cap + 1);
        return (int64_t)(g.cap + 1);
    }
    uint64_t s = home_of(g, h);
    for (uint64_t p = 0; p <= g.sub_mask; p++) {
        uint64_t* kp = g.keys + s;
        uint64_t cur = g_load_volatile(kp);  // a stale kEmpty only costs a CAS
        if (cur == key) return (int64_t)s;
        if (cur == kEmpty) {
            unsigned long long old = g_atomic_cas((unsigned long long*)kp, (unsigned long long)kEmpty,
                                                  (unsigned long long)key);
            if (old == kEmpty) {
                *is_new = true;
                return (int64_t)s;
            }
            if (old == key) return (int64_t)s;
        }
        s = next_in_sub(g, s);
    }
    return -1;
}

// As g_slot1 for a key that is usually absent (the merge of a workgroup's table into its own
// sub-table): every probe is one CAS instead of a read and then a CAS.
__device__ __forceinline__ int64_t g_slot1_claim(const GView& g, uint64_t key, uint64_t h, bool* is_new) {
    *is_new = false;
    uint64_t s = home_of(g, h);
    for (uint64_t p = 0; p <= g.sub_mask; p++) {
        const unsigned long long old = g_atomic_cas((unsigned long long*)(g.keys + s), (unsigned long long)kEmpty,
                                                    (unsigned long long)key);
        if (old == kEmpty) {
            *is_new = true;
            return (int64_t)s;
        }
        if (old == key) return (int64_t)s;
        s = next_in_sub(g, s);
    }
    return -1;
}

// Slot of a multi-word key: the last word is claimed (0 -> kBusy), the other words written,
// then the last word published with release order; readers that see kBusy wait for it.
template <int KW>
__device__ __forceinline__ int64_t g_slotN(const GView& g, const uint64_t (&k)[KW], uint64_t h,
                                           bool* is_new) {
    *is_new = false;
    uint64_t s = home_of(g, h);
    for (uint64_t p = 0; p <= g.sub_mask; p++) {
        uint64_t* kp = g.keys + s * KW;
        uint64_t last = __hip_atomic_load(kp + (KW - 1), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (last == 0ull) {
            unsigned long long old = atomicCAS((unsigned long long*)(kp + (KW - 1)), 0ull,
                                               (unsigned long long)kBusy);
            if (old == 0ull) {
#pragma unroll
                for (int w = 0; w < KW - 1; w++) st_relaxed(kp + w, k[w]);
                __hip_atomic_store(kp + (KW - 1), k[KW - 1], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                *is_new = true;
                return (int64_t)s;
            }
            last = old;
        }
        while (last == kBusy) {
            __builtin_amdgcn_s_sleep(1);
            last = __hip_atomic_load(kp + (KW - 1), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (last == k[KW - 1]) {
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW - 1; w++) eq = eq && (ld_relaxed(kp + w) == k[w]);
            if (eq) return (int64_t)s;
        }
        s = next_in_sub(g, s);
    }
    return -1;
}

// Whether the table holds a key (no concurrent writers: a plain probe).  The NULL key and the
// key equal to kEmpty of a single-word table live in their side slots.
__device__ __forceinline__ bool g_has1(const GView& g, uint64_t key, bool knull, uint64_t h) {
    if (knull) return g.keys[g.cap] != 0ull;
    if (key == kEmpty) return g.keys[g.cap + 1] != 0ull;
    uint64_t s = home_of(g, h);
    for (uint64_t p = 0; p <= g.sub_mask; p++) {
        const uint64_t cur = g.keys[s];
        if (cur == key) return true;
        if (cur == kEmpty) return false;
        s = next_in_sub(g, s);
    }
    return false;
}

template <int KW>
__device__ __forceinline__ bool g_hasN(const GView& g, const uint64_t (&k)[KW], uint64_t h) {
    uint64_t s = home_of(g, h);
    for (uint64_t p = 0; p <= g.sub_mask; p++) {
        const uint64_t* kp = g.keys + s * KW;
        const uint64_t last = kp[KW - 1];
        if (last == 0ull) return false;
        bool eq = last == k[KW - 1];
#pragma unroll
        for (int w = 0; w < KW - 1; w++) eq = eq && kp[w] == k[w];
        if (eq) return true;
        s = next_in_sub(g, s);
    }
    return false;
}

// Atomic merge of a partial into slot s (concurrent writers allowed).
__device__ __forceinline__ void g_accum_atomic(const GView& g, int64_t s, const Part& p, uint32_t fields) {
    if (p.cnt) g_atomic_add((unsigned long long*)(g.cnt + s), (unsigned long long)p.cnt);
    if ((fields & F_CNTALL) && p.cntall) g_atomic_add((unsigned long long*)(g.cntall + s), (unsigned long long)p.cntall);
    if (p.cnt) {
        if (fields & F_SUM) {
            if (g.real_sum)
                g_atomic_add((double*)(g.sum + s), __longlong_as_double((long long)p.sum));
            else
                g_atomic_add((unsigned long long*)(g.sum + s), (unsigned long long)p.sum);
        }
        // the read first skips atomics that cannot change the value
        if ((fields & F_MIN) && p.mn < g_load_volatile(g.mn + s))
            g_atomic_min((unsigned long long*)(g.mn + s), (unsigned long long)p.mn);
        if ((fields & F_MAX) && p.mx > g_load_volatile(g.mx + s))
            g_atomic_max((unsigned long long*)(g.mx + s), (unsigned long long)p.mx);
    }
}

// Plain merge of a partial into slot s: only for a slot no other thread touches meanwhile.
__device__ __forceinline__ void g_accum_owned(const GView& g, int64_t s, const Part& p, uint32_t fields) {
    if (p.cnt) g.cnt[s] += p.cnt;
    if ((fields & F_CNTALL) && p.cntall) g.cntall[s] += p.cntall;
    if (p.cnt) {
        if (fields & F_SUM) {
            if (g.real_sum)
                g.sum[s] = (unsigned long long)__double_as_longlong(
                    __longlong_as_double((long long)g.sum[s]) + __longlong_as_double((long long)p.sum));
            else
                g.sum[s] += p.sum;
        }
        if ((fields & F_MIN) && p.mn < g.mn[s]) g.mn[s] = p.mn;
        if ((fields & F_MAX) && p.mx > g.mx[s]) g.mx[s] = p.mx;
    }
}

// The states of a slot this workgroup has just claimed (still at their initial values, and no
// other workgroup writes its sub-table): stored without reading them first.
__device__ __forceinline__ void g_store_new(const GView& g, int64_t s, const Part& p, uint32_t fields) {
    g.cnt[s] = p.cnt;
    if (fields & F_CNTALL) g.cntall[s] = p.cntall;
    if (fields & F_SUM) g.sum[s] = p.cnt ? p.sum : 0ull;
    if (fields & F_MIN) g.mn[s] = p.cnt ? p.mn : ~0ull;
    if (fields & F_MAX) g.mx[s] = p.cnt ? p.mx : 0ull;
}

// Count `n` new keys of one workgroup; reports a full table against the logical capacity.
__device__ __forceinline__ void g_count_new(const GView& g, uint64_t n) {
    if (!n) return;
    unsigned long long c = g_atomic_add(g.nkeys, (unsigned long long)n) + n;
    if (c > g.logical_cap) g_set_error(g, QK_HASH_TABLE_FULL);
}

}  // namespace qk
