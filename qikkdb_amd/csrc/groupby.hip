// groupby.hip - hash GROUP BY with a fused WHERE, for gfx950.
//
// Counterpart of GPUGroupBy<AGG,O,K,V> (QueryEngine/GPUCore/GPUGroupBy.cuh:68-791) and its
// composite-key specialisation (GPUGroupByMultiKey.cuh:156-1325).  Design differences:
//  * One table keeps several aggregates per group (the reference keeps one, S11).
//  * Single-word keys (one key column, or composite keys packed into 64 bits) go through the
//    two-phase partitioned path (twophase.hip): heavy hitters aggregate in LDS, every other
//    row is partitioned without atomics and merged by the partition's owning workgroup.  The
//    reference instead spreads each slot over 256 replicas in HBM (IGroupBy.h:5,
//    GPUGroupBy.cuh:181-191) and folds them in GetResults.
//  * Multi-word keys and small blocks use a one-pass kernel: a per-workgroup LDS table in
//    front of the global table, merged with atomics at the end.
//  * The global table is a structure of arrays: key words, counts, sums, minima, maxima.
//    Single-word keys are claimed with a 64-bit CAS; multi-word keys by claiming the last word
//    (empty -> busy -> key).  MIN/MAX of every type use native unsigned 64-bit atomics on an
//    order-preserving encoding; SUM of FLOAT/DOUBLE accumulates in FP64, integer SUM in
//    wrapping int64.
//  * Results come out in slot order through a look-back compaction, in one kernel.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "groupby.hpp"
#include "lookback.hpp"
#include "twophase.hpp"

namespace qk {

namespace {

constexpr int kSteps = 4;        // rows per thread per tile
constexpr uint64_t kMaxLogical = (uint64_t)1 << 30;  // GB_BUFFER_SIZE_MAX (IGroupBy.h:6)
constexpr int kLdsProbes = 8;    // LDS probe limit before a row goes to the global table
constexpr int kEmitBlock = 256;
constexpr int kEmitSteps = 16;
constexpr int kEmitTile = kEmitBlock * kEmitSteps;

// ---- LDS table of the one-pass kernel ----------------------------------------------------
// Arrays of H hashed slots + 2 side slots (NULL key, key == kEmpty).
struct Lds {
    uint64_t* key;
    uint64_t* sum;
    uint64_t* mn;
    uint64_t* mx;
    uint32_t* cnt;
    uint32_t* cntall;
    int32_t H;
};

__host__ __device__ __forceinline__ int lds_slot_bytes(uint32_t fields) {
    return 8 + 4 + ((fields & F_CNTALL) ? 4 : 0) + ((fields & F_SUM) ? 8 : 0) +
           ((fields & F_MIN) ? 8 : 0) + ((fields & F_MAX) ? 8 : 0);
}

__device__ __forceinline__ Lds lds_carve(void* base, int32_t H, uint32_t fields) {
    Lds t;
    const int T = H + 2;
    char* p = (char*)base;
    t.key = (uint64_t*)p; p += (size_t)T * 8;
    t.sum = (fields & F_SUM) ? (uint64_t*)p : nullptr; p += (fields & F_SUM) ? (size_t)T * 8 : 0;
    t.mn = (fields & F_MIN) ? (uint64_t*)p : nullptr; p += (fields & F_MIN) ? (size_t)T * 8 : 0;
    t.mx = (fields & F_MAX) ? (uint64_t*)p : nullptr; p += (fields & F_MAX) ? (size_t)T * 8 : 0;
    t.cnt = (uint32_t*)p; p += (size_t)T * 4;
    t.cntall = (fields & F_CNTALL) ? (uint32_t*)p : nullptr;
    t.H = H;
    return t;
}

__device__ __forceinline__ int lds_find(const Lds& t, uint64_t key, bool knull, uint64_t h) {
    if (knull) {
        if (t.key[t.H] == 0ull) t.key[t.H] = 1ull;
        return t.H;
    }
    if (key == kEmpty) {
        if (t.key[t.H + 1] == 0ull) t.key[t.H + 1] = 1ull;
        return t.H + 1;
    }
    uint32_t s = (uint32_t)(((h >> 32) * (uint64_t)(uint32_t)t.H) >> 32);
#pragma unroll 1
    for (int p = 0; p < kLdsProbes; p++) {
        uint64_t cur = __hip_atomic_load(&t.key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == key) return (int)s;
        if (cur == kEmpty) {
            unsigned long long old = atomicCAS((unsigned long long*)&t.key[s], (unsigned long long)kEmpty,
                                               (unsigned long long)key);
            if (old == kEmpty || old == key) return (int)s;
        }
        s = (s + 1 == (uint32_t)t.H) ? 0u : s + 1;
    }
    return -1;
}

template <bool REAL>
__device__ __forceinline__ void lds_accum(const Lds& t, int s, const Part& p, uint32_t fields) {
    if (p.cnt) atomicAdd(&t.cnt[s], (uint32_t)p.cnt);
    if (fields & F_CNTALL) atomicAdd(&t.cntall[s], (uint32_t)p.cntall);
    if (p.cnt) {
        if (fields & F_SUM) {
            if (REAL)
                atomicAdd((double*)&t.sum[s], __longlong_as_double((long long)p.sum));
            else
                atomicAdd((unsigned long long*)&t.sum[s], (unsigned long long)p.sum);
        }
        if (fields & F_MIN) atomicMin((unsigned long long*)&t.mn[s], (unsigned long long)p.mn);
        if (fields & F_MAX) atomicMax((unsigned long long*)&t.mx[s], (unsigned long long)p.mx);
    }
}

// ---- the one-pass row kernel --------------------------------------------------------------
// MODE 0: every row; 1: rows whose int8 mask is non-zero; 2: rows passing the predicates.
// Row r of tile T is T*(kSteps*blockDim) + k*blockDim + threadIdx.x, so each step's loads
// are coalesced spans.  Workgroups are persistent (grid-stride over tiles) so that the LDS
// table sees many rows before it is merged.
template <int KW, typename VT, int MODE>
__global__ void __launch_bounds__(1024) k_gb_rows(GView g, KeySpec ks, KeyCols kc,
                                                  const VT* __restrict__ vals,
                                                  const uint64_t* __restrict__ vnull,
                                                  const int8_t* __restrict__ mask, PredSet preds,
                                                  uint32_t vpred, int64_t n, uint32_t fields,
                                                  int32_t H) {
    extern __shared__ uint64_t lds_raw[];
    __shared__ unsigned int s_new;
    constexpr bool HAS_V = VTraits<VT>::has;
    constexpr bool REAL = VTraits<VT>::real;
    const bool use_lds = (KW == 1) && H > 0;
    Lds t{};
    if (threadIdx.x == 0) s_new = 0;
    if (use_lds) {
        t = lds_carve(lds_raw, H, fields);
        for (int s = threadIdx.x; s < H + 2; s += blockDim.x) {
            t.key[s] = s < H ? kEmpty : 0ull;
            t.cnt[s] = 0;
            if (t.cntall) t.cntall[s] = 0;
            if (t.sum) t.sum[s] = 0;
            if (t.mn) t.mn[s] = ~0ull;
            if (t.mx) t.mx[s] = 0ull;
        }
    }
    __syncthreads();

    const int64_t bd = blockDim.x;
    const int64_t tile_rows = bd * kSteps;
    const int64_t ntiles = (n + tile_rows - 1) / tile_rows;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t key[kSteps][KW];
        bool kn[kSteps], act[kSteps], vn[kSteps];
        typename std::conditional<HAS_V, VT, int>::type val[kSteps];
#pragma unroll
        for (int k = 0; k < kSteps; k++) {
            const int64_t r = tile * tile_rows + k * bd + threadIdx.x;
            act[k] = r < n;
            vn[k] = false;
            kn[k] = false;
            val[k] = {};
            if (act[k]) {
                load_key<KW>(ks, kc, r, key[k], kn[k]);
                if constexpr (HAS_V) val[k] = vals[r];
                vn[k] = vnull && null_bit(vnull, r);
                if (MODE == 1) act[k] = mask[r] != 0;
                if (MODE == 4)
                    act[k] = !(preds.rnull && null_bit(preds.rnull, r)) && col_in_range(preds, range_col(preds, r));
            }
        }
        if constexpr (MODE == 3 && HAS_V) {
#pragma unroll
            for (int k = 0; k < kSteps; k++) act[k] = act[k] && !vn[k] && in_range<VT>(preds.vr, val[k]);
        }
        if (MODE == 2) {
#pragma unroll
            for (int k = 0; k < kSteps; k++) {
                if (!act[k]) continue;
                const int64_t r = tile * tile_rows + k * bd + threadIdx.x;
                bool pass = true;
                for (int p = 0; p < preds.n && pass; p++) {
                    const qk_predicate& q = preds.p[p];
                    if constexpr (HAS_V) {
                        if ((vpred >> p) & 1u) {
                            pass = !vn[k] && pred_on_value(q, val[k], r);
                            continue;
                        }
                    }
                    pass = eval_pred(q, r);
                }
                act[k] = pass;
            }
        }
#pragma unroll
        for (int k = 0; k < kSteps; k++) {
            if (!act[k]) continue;
            Part p;
            if constexpr (HAS_V) p = row_part<VT>(val[k], vn[k]); else p = row_part<NoValue>(NoValue{}, vn[k]);
            const uint64_t h = key_hash<KW>(key[k]);
            bool is_new = false;
            if constexpr (KW == 1) {
                int ls = use_lds ? lds_find(t, key[k][0], kn[k], h) : -1;
                if (ls >= 0) {
                    lds_accum<REAL>(t, ls, p, fields);
                } else {
                    const int64_t gs = g_slot1(g, key[k][0], kn[k], h, &is_new);
                    if (gs >= 0) g_accum_atomic(g, gs, p, fields); else g_set_lost(g);
                }
            } else {
                const int64_t gs = g_slotN<KW>(g, key[k], h, &is_new);
                if (gs >= 0) g_accum_atomic(g, gs, p, fields); else g_set_lost(g);
            }
            if (is_new) atomicAdd(&s_new, 1u);
        }
    }

    if (use_lds) {
        __syncthreads();
        for (int s = threadIdx.x; s < H + 2; s += blockDim.x) {
            const uint64_t k = t.key[s];
            if (s < H ? (k == kEmpty) : (k == 0ull)) continue;
            Part p;
            p.cnt = t.cnt[s];
            p.cntall = t.cntall ? t.cntall[s] : 0;
            p.sum = t.sum ? t.sum[s] : 0;
            p.mn = t.mn ? t.mn[s] : ~0ull;
            p.mx = t.mx ? t.mx[s] : 0ull;
            bool is_new = false;
            int64_t gs;
            if (s == H) {
                gs = g_slot1(g, 0ull, true, 0, &is_new);
            } else if (s == H + 1) {
                gs = g_slot1(g, kEmpty, false, 0, &is_new);
            } else {
                gs = g_slot1(g, k, false, hash64(k), &is_new);
            }
            if (is_new) atomicAdd(&s_new, 1u);
            if (gs >= 0) g_accum_atomic(g, gs, p, fields); else g_set_lost(g);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) g_count_new(g, s_new);
}

// ---- table initialisation ----------------------------------------------------------------
__global__ void k_gb_init(GView g, uint32_t fields) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the table's counters (TableCounters)
        *g.nkeys = 0;
        *g.err = 0;
        *g.lost = 0;
    }
    const uint64_t total = g.cap + 2;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += stride) {
        for (int w = 0; w < g.kw; w++) g.keys[s * g.kw + w] = (g.kw == 1 && s < g.cap) ? kEmpty : 0ull;
        g.cnt[s] = 0;
        if (g.cntall) g.cntall[s] = 0;
        if (g.sum) g.sum[s] = 0;
        if (g.mn) g.mn[s] = ~0ull;
        if (g.mx) g.mx[s] = 0;
    }
}

// Clears n words in each of up to kMaxKeys + kMaxAggs buffers (blockIdx.y picks the buffer).
struct ZeroList {
    int32_t n;
    uint64_t* p[kMaxKeys + kMaxAggs];
};
__global__ void k_zero_words(ZeroList z, int64_t n) {
    uint64_t* p = z.p[blockIdx.y];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---- results ------------------------------------------------------------------------------
struct OutSpec {
    int32_t naggs;
    int32_t vtype;
    int32_t agg[kMaxAggs];
    int32_t otype[kMaxAggs];
    void* out[kMaxAggs];
    uint64_t* onull[kMaxAggs];
    void* kout[kMaxKeys];
    uint64_t* konull[kMaxKeys];
};

__device__ __forceinline__ bool slot_used(const GView& g, uint64_t s) {
    const uint64_t* k = g.keys + s * (uint64_t)g.kw;
    if (g.kw == 1) return s < g.cap ? (k[0] != kEmpty) : (k[0] != 0ull);
    return s < g.cap && k[g.kw - 1] != 0ull;
}

__device__ __forceinline__ uint64_t emit_slot(const GView& g, uint64_t i) {
    return i == 0 ? g.cap : (i <= g.cap ? i - 1 : i);
}

__device__ __forceinline__ void put_typed(void* out, int type, int64_t pos, int64_t iv, double dv,
                                          bool real) {
    switch (type) {
    case QK_INT8: ((int8_t*)out)[pos] = (int8_t)(real ? (int64_t)dv : iv); break;
    case QK_INT: ((int32_t*)out)[pos] = (int32_t)(real ? (int64_t)dv : iv); break;
    case QK_LONG: ((int64_t*)out)[pos] = real ? (int64_t)dv : iv; break;
    case QK_FLOAT: ((float*)out)[pos] = real ? (float)dv : (float)iv; break;
    default: ((double*)out)[pos] = real ? dv : (double)iv; break;
    }
}

__device__ __forceinline__ void set_bit(uint64_t* m, int64_t pos) {
    atomicOr((unsigned long long*)(m + (pos >> 6)), 1ull << (pos & 63));
}

// Key columns of the group in slot s.
__device__ __forceinline__ void emit_key(const GView& g, const KeySpec& ks, const OutSpec& os,
                                         uint64_t s, int64_t pos) {
    const uint64_t* r = g.keys + s * (uint64_t)g.kw;
    for (int c = 0; c < ks.nkeys; c++) {
        const int type = ks.type[c];
        bool isnull = false;
        uint64_t v;
        if (ks.kw == 1) {
            const uint64_t w = (s == g.cap) ? 0ull : (s == g.cap + 1 ? kEmpty : r[0]);
            if (ks.single) {
                isnull = (s == g.cap);
                v = w;
            } else {
                v = w >> ks.shift[c];
            }
        } else {
            isnull = ks.nullable && ((r[ks.nword[c]] >> ks.nshift[c]) & 1ull);
            v = r[ks.word[c]] >> ks.shift[c];
        }
        if (isnull) v = 0;
        void* o = os.kout[c];
        if (o) {
            switch (type) {
            case QK_INT: ((int32_t*)o)[pos] = (int32_t)(uint32_t)v; break;
            case QK_LONG: ((int64_t*)o)[pos] = (int64_t)v; break;
            case QK_FLOAT: ((uint32_t*)o)[pos] = (uint32_t)v; break;
            case QK_DOUBLE: ((uint64_t*)o)[pos] = v; break;
            default: ((int8_t*)o)[pos] = (int8_t)(uint32_t)v; break;
            }
        }
        if (isnull && os.konull[c]) set_bit(os.konull[c], pos);
    }
}

__device__ __forceinline__ void emit_aggs(const GView& g, const OutSpec& os, uint64_t s, int64_t pos) {
    const uint64_t cnt = g.cnt[s];
    const int vt = os.vtype;
    const bool vreal = vt == QK_FLOAT || vt == QK_DOUBLE;
    for (int a = 0; a < os.naggs; a++) {
        const int ag = os.agg[a];
        void* o = os.out[a];
        if (ag == QK_AGG_NONE || !o) continue;
        const int ot = os.otype[a];
        bool isnull = false;
        switch (ag) {
        case QK_AGG_COUNT: ((int64_t*)o)[pos] = (int64_t)cnt; break;
        case QK_AGG_COUNT_ALL: ((int64_t*)o)[pos] = (int64_t)g.cntall[s]; break;
        case QK_AGG_SUM: {
            isnull = cnt == 0;
            if (vreal) put_typed(o, ot, pos, 0, isnull ? 0.0 : __longlong_as_double((long long)g.sum[s]), true);
            else put_typed(o, ot, pos, isnull ? 0 : (int64_t)g.sum[s], 0.0, false);
            break;
        }
        case QK_AGG_MIN:
        case QK_AGG_MAX: {
            isnull = cnt == 0;
            const uint64_t e = ag == QK_AGG_MIN ? g.mn[s] : g.mx[s];
            if (vreal) {
                double d = isnull ? (ag == QK_AGG_MIN ? (vt == QK_FLOAT ? 3.4028234663852886e38 : 1.7976931348623157e308)
                                                      : (vt == QK_FLOAT ? -3.4028234663852886e38 : -1.7976931348623157e308))
                                  : dec_f64(e);
                put_typed(o, ot, pos, 0, d, true);
            } else {
                int64_t lim_hi = vt == QK_INT ? INT32_MAX : (vt == QK_INT8 ? INT8_MAX : INT64_MAX);
                int64_t lim_lo = vt == QK_INT ? INT32_MIN : (vt == QK_INT8 ? INT8_MIN : INT64_MIN);
                int64_t v = isnull ? (ag == QK_AGG_MIN ? lim_hi : lim_lo) : dec_i64(e);
                put_typed(o, ot, pos, v, 0.0, false);
            }
            break;
        }
        case QK_AGG_AVG: {
            // sum / count with the sum in the value type (GPUGroupBy.cuh:536-558,
            // GPUBinary.cuh:126-143); count 0 gives the NULL constant (NullConstants.cuh:8-26)
            isnull = cnt == 0;
            if (vreal) {
                double d = isnull ? __longlong_as_double(0x7ff8000000000000ll)
                                  : __longlong_as_double((long long)g.sum[s]) / (double)cnt;
                put_typed(o, ot, pos, 0, d, true);
            } else {
                int64_t v;
                const uint64_t sm = g.sum[s];
                if (vt == QK_INT) v = isnull ? INT32_MIN : (int64_t)(int32_t)((int64_t)(int32_t)(uint32_t)sm / (int64_t)cnt);
                else if (vt == QK_INT8) v = isnull ? INT8_MIN : (int64_t)(int8_t)((int64_t)(int8_t)(uint8_t)sm / (int64_t)cnt);
                else v = isnull ? INT64_MIN : (int64_t)sm / (int64_t)cnt;
                put_typed(o, ot, pos, v, 0.0, false);
            }
            break;
        }
        default: break;
        }
        if (isnull && os.onull[a]) set_bit(os.onull[a], pos);
    }
}

// Row order: the NULL-key group first (the reference keeps it in slot 0, and its callers read
// "first key is NULL" from bit 0 of the key null mask, GPUGroupBy.cuh:727-733), then the hashed
// slots in slot order, then the side slot of the key equal to the empty marker.  Position i of
// that order is slot emit_slot(i).  Slots whose bit is set in `skip` (the seeded keys, emitted
// by k_gb_emit_seeded) are left out; the other groups go to rows base, base + 1, ...
__global__ void __launch_bounds__(kEmitBlock) k_gb_emit(GView g, KeySpec ks, OutSpec os,
                                                        uint64_t* status, uint32_t* ticket,
                                                        uint32_t epoch, int64_t tiles,
                                                        int32_t* total, int64_t limit,
                                                        const uint64_t* __restrict__ skip,
                                                        int64_t base) {
    constexpr int W = kEmitBlock / 64;
    __shared__ int64_t s_tile;
    __shared__ int32_t s_cnt[kEmitSteps][W];
    __shared__ uint32_t s_excl;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_tile = lb_ticket(ticket, tiles);
    __syncthreads();
    const int64_t tile = s_tile;
    const uint64_t nslots = g.cap + 2;
    const uint64_t lt_mask = (1ull << lane) - 1;
    uint32_t keep = 0, pre[kEmitSteps];
#pragma unroll
    for (int k = 0; k < kEmitSteps; k++) {
        const uint64_t i = (uint64_t)tile * kEmitTile + (uint64_t)k * kEmitBlock + tid;
        const uint64_t s = emit_slot(g, i);
        const bool sel = i < nslots && slot_used(g, s) && !(skip && ((skip[s >> 6] >> (s & 63)) & 1ull));
        const uint64_t b = __ballot(sel);
        pre[k] = (uint32_t)__popcll(b & lt_mask);
        if (lane == 0) s_cnt[k][wid] = (int32_t)__popcll(b);
        keep |= (sel ? 1u : 0u) << k;
    }
    __syncthreads();
    int32_t agg = 0;
    for (int k = 0; k < kEmitSteps; k++)
        for (int w = 0; w < W; w++) agg += s_cnt[k][w];
    if (wid == 0) {
        uint32_t e = lb_exclusive(status, epoch, tile, (uint32_t)agg);
        if (lane == 0) {
            s_excl = e;
            if (tile == tiles - 1) *total = (int32_t)(e + (uint32_t)agg);
        }
    }
    __syncthreads();
    int64_t step_base = base + (int64_t)s_excl;
    for (int k = 0; k < kEmitSteps; k++) {
        int64_t wave_base = step_base;
        for (int w = 0; w < wid; w++) wave_base += s_cnt[k][w];
        if ((keep >> k) & 1u) {
            const uint64_t s = emit_slot(g, (uint64_t)tile * kEmitTile + (uint64_t)k * kEmitBlock + tid);
            const int64_t pos = wave_base + pre[k];
            if (pos < limit) {  // outputs hold `limit` rows; the host checks the total
                emit_key(g, ks, os, s, pos);
                emit_aggs(g, os, s, pos);
            }
        }
        int32_t st = 0;
        for (int w = 0; w < W; w++) st += s_cnt[k][w];
        step_base += st;
    }
}

// Row i of the results is seed key i (the pre-seeded table keeps its keys' order).
__global__ void k_gb_emit_seeded(GView g, KeySpec ks, OutSpec os, const int64_t* __restrict__ seed_slot,
                                 int32_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t s = (uint64_t)seed_slot[i];
        emit_key(g, ks, os, s, i);
        emit_aggs(g, os, s, i);
    }
}

// Insert n keys of one key column (key 0 is the NULL group when first_null) as groups of
// their own, whatever rows later reach them; slot_out[i] receives key i's slot and the slot's
// bit is set in `bits`.
__global__ void k_gb_seed(GView g, int32_t ktype, const void* __restrict__ keys, int32_t n,
                          int32_t first_null, int64_t* slot_out, uint64_t* bits) {
    __shared__ unsigned int s_new;
    if (threadIdx.x == 0) s_new = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool kn = first_null && i == 0;
        const uint64_t k = kn ? 0ull : key_single(ktype, keys, i);
        bool is_new = false;
        const int64_t s = g_slot1(g, k, kn, hash64(k), &is_new);
        if (s < 0) {
            g_set_error(g, QK_HASH_TABLE_FULL);
        } else {
            atomicOr((unsigned long long*)(bits + (s >> 6)), 1ull << (s & 63));
        }
        if (is_new) atomicAdd(&s_new, 1u);
        slot_out[i] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) g_count_new(g, s_new);
}

// Fold every group of table `src` into table `dst` (same key layout and aggregate states):
// each group's key is inserted into dst and its states merged with atomics, so the result is
// that of one table that saw both tables' rows.
template <int KW>
__global__ void __launch_bounds__(256) k_gb_merge(GView dst, GView src, uint32_t fields) {
    __shared__ unsigned int s_new;
    if (threadIdx.x == 0) s_new = 0;
    __syncthreads();
    const uint64_t nslots = src.cap + 2;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += (uint64_t)gridDim.x * blockDim.x) {
        if (!slot_used(src, s)) continue;
        Part p;
        p.cnt = src.cnt[s];
        p.cntall = src.cntall ? src.cntall[s] : 0ull;
        p.sum = src.sum ? src.sum[s] : 0ull;
        p.mn = src.mn ? src.mn[s] : ~0ull;
        p.mx = src.mx ? src.mx[s] : 0ull;
        bool is_new = false;
        int64_t d;
        if constexpr (KW == 1) {
            if (s == src.cap) {
                d = g_slot1(dst, 0ull, true, 0, &is_new);
            } else {
                const uint64_t k = s == src.cap + 1 ? kEmpty : src.keys[s];
                d = g_slot1(dst, k, false, hash64(k), &is_new);
            }
        } else {
            uint64_t k[KW];
#pragma unroll
            for (int w = 0; w < KW; w++) k[w] = src.keys[s * KW + w];
            d = g_slotN<KW>(dst, k, key_hash<KW>(k), &is_new);
        }
        if (d < 0) {
            g_set_lost(dst);
            continue;
        }
        if (is_new) atomicAdd(&s_new, 1u);
        g_accum_atomic(dst, d, p, fields);
    }
    __syncthreads();
    if (threadIdx.x == 0) g_count_new(dst, s_new);
}

// Rows of a block still to aggregate after the table dropped partials (a growing table): rows
// that pass the block's WHERE (MODE 0 none, 1 int8 mask, 2 predicate list) and whose key the
// table does not hold.  out[r] = 1 for those rows, 0 otherwise.  `out` may be `mask` itself
// (each thread reads its row's mask entry before writing it).
template <int KW, int MODE>
__global__ void __launch_bounds__(256) k_gb_absent(GView g, KeySpec ks, KeyCols kc,
                                                   const int8_t* mask, PredSet preds, int64_t n,
                                                   int8_t* out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        bool pass = true;
        if (MODE == 1) pass = mask[r] != 0;
        if (MODE == 2)
            for (int p = 0; p < preds.n && pass; p++) pass = eval_pred(preds.p[p], r);
        bool absent = false;
        if (pass) {
            uint64_t k[KW];
            bool kn = false;
            load_key<KW>(ks, kc, r, k, kn);
            if constexpr (KW == 1)
                absent = !g_has1(g, k[0], kn, hash64(k[0]));
            else
                absent = !g_hasN<KW>(g, k, key_hash<KW>(k));
        }
        out[r] = (int8_t)(pass && absent);
    }
}

// ---- launch helpers -------------------------------------------------------------------------

template <int KW, typename VT, int MODE>
int launch_rows_t(const GView& g, const KeySpec& ks, const KeyCols& kc, const void* vals,
                  const uint64_t* vnull, const int8_t* mask, const PredSet& ps, uint32_t vpred,
                  int64_t n, uint32_t fields, int32_t H, int block, int wgs_per_cu,
                  hipStream_t s) {
    auto kern = k_gb_rows<KW, VT, MODE>;
    // gfx950 launches up to 160 KB of LDS per workgroup without a function attribute
    const size_t lds = (KW == 1 && H > 0) ? (size_t)(H + 2) * lds_slot_bytes(fields) : 0;
    int occ = 0;
    QK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, block, lds));
    if (occ < 1) return fail(QK_ERROR, "group by: launch shape (block %d, LDS %zu B) does not fit", block, lds);
    if (wgs_per_cu > 0 && wgs_per_cu < occ) occ = wgs_per_cu;
    const int64_t tile_rows = (int64_t)block * kSteps;
    const int64_t ntiles = (n + tile_rows - 1) / tile_rows;
    int64_t grid = (int64_t)device_info().cus * occ;
    if (grid > ntiles) grid = ntiles;
    if (grid < 1) grid = 1;
    kern<<<(unsigned)grid, block, lds, s>>>(g, ks, kc, (const VT*)vals, vnull, mask, ps, vpred, n,
                                            fields, (KW == 1) ? H : 0);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

template <int KW, typename VT>
int launch_rows_m(int mode, const GView& g, const KeySpec& ks, const KeyCols& kc, const void* vals,
                  const uint64_t* vnull, const int8_t* mask, const PredSet& ps, uint32_t vpred,
                  int64_t n, uint32_t fields, int32_t H, int block, int wpc, hipStream_t s) {
    switch (mode) {
    case 0: return launch_rows_t<KW, VT, 0>(g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case 1: return launch_rows_t<KW, VT, 1>(g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case 2: return launch_rows_t<KW, VT, 2>(g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case 4: return launch_rows_t<KW, VT, 4>(g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    default:
        if constexpr (VTraits<VT>::has)
            return launch_rows_t<KW, VT, 3>(g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
        return fail(QK_ERROR, "value range filter without a value column");
    }
}

template <int KW>
int launch_rows_v(int vtype, int mode, const GView& g, const KeySpec& ks, const KeyCols& kc,
                  const void* vals, const uint64_t* vnull, const int8_t* mask, const PredSet& ps,
                  uint32_t vpred, int64_t n, uint32_t fields, int32_t H, int block, int wpc,
                  hipStream_t s) {
    if (!vals) return launch_rows_m<KW, NoValue>(mode, g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    switch (vtype) {
    case QK_INT: return launch_rows_m<KW, int32_t>(mode, g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case QK_LONG: return launch_rows_m<KW, int64_t>(mode, g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case QK_FLOAT: return launch_rows_m<KW, float>(mode, g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    case QK_DOUBLE: return launch_rows_m<KW, double>(mode, g, ks, kc, vals, vnull, mask, ps, vpred, n, fields, H, block, wpc, s);
    default: return fail(QK_ERROR, "group by: unsupported value type %d", vtype);
    }
}

}  // namespace
}  // namespace qk

using namespace qk;

// ---- the table object ----------------------------------------------------------------------
struct qk_groupby_s {
    int device = 0;
    int32_t nkeys = 0, key_types[kMaxKeys] = {}, value_type = 0, naggs = 0;
    int32_t aggs[kMaxAggs] = {}, out_types[kMaxAggs] = {};
    uint32_t flags = 0, fields = 0;
    KeySpec ks{};
    uint64_t logical_cap = 0;
    int32_t reported_max = 0;
    uint64_t cap = 0;
    int32_t parts = 1;      // sub-tables of the hashed slots
    int32_t rec_parts = 1;  // partitions of the two-phase path's cold rows (divides parts)
    bool real_sum = false;
    // storage (structure of arrays)
    uint64_t* keys = nullptr;
    unsigned long long *cnt = nullptr, *cntall = nullptr, *sum = nullptr, *mn = nullptr, *mx = nullptr;
    TableCounters* ctr = nullptr;
    // two-phase path
    TwoPhaseBufs tp;
    int64_t twophase_min_rows = (int64_t)1 << 20;
    int32_t twophase_mode = -1;  // -1 automatic, 0 never, 1 always (single-word keys)
    // launch tuning of the one-pass kernel
    int32_t block = 512;
    int32_t lds_budget = -1;  // bytes of LDS per workgroup; -1 = half of a CU's LDS, 0 = none
    int32_t wgs_per_cu = 0;   // 0 = as many as fit
    int32_t last_H = 0;
    int32_t last_path = 0;
    int64_t grows = 0;      // growths of a QK_GB_AUTO_GROW table
    // pre-seeded keys (GPUGroupBy(maxHashCount, keys, firstKeyIsNull)): their slots in seed
    // order and a bit per slot; GetResults emits them first, in that order
    int64_t* seed_slot = nullptr;
    uint64_t* seed_bits = nullptr;
    int32_t seed_n = 0;
    // profiling
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t events_used = 0;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
};

namespace qk {
namespace {

GView view_of(const qk_groupby_s* h) {
    GView g{};  // zeroed padding: views are compared bytewise
    g.keys = h->keys;
    g.cnt = h->cnt;
    g.cntall = h->cntall;
    g.sum = h->sum;
    g.mn = h->mn;
    g.mx = h->mx;
    g.cap = h->cap;
    int pbits = 0;
    while ((1 << pbits) < h->parts) pbits++;
    int cbits = 0;
    while (((uint64_t)1 << cbits) < h->cap) cbits++;
    g.part_shift = pbits == 0 ? 64 : 64 - pbits;
    g.sub_bits = cbits - pbits;
    g.sub_mask = ((uint64_t)1 << g.sub_bits) - 1;
    g.kw = h->ks.kw;
    g.real_sum = h->real_sum ? 1 : 0;
    g.nkeys = &h->ctr->nkeys;
    g.err = &h->ctr->err;
    g.logical_cap = h->logical_cap;
    g.lost = &h->ctr->lost;
    return g;
}

// Slot count and sub-table layout of a table holding `logical` keys (before it reports itself
// full): 2 slots per key, rounded up to a power of two.
void layout_table(qk_groupby_s* h, uint64_t logical) {
    h->logical_cap = logical;
    h->reported_max = (int32_t)(h->nkeys == 1 ? logical + 1 : logical);  // + 1 for the NULL key
    uint64_t cap = 64;
    while (cap < 2 * logical) cap <<= 1;
    h->cap = cap;
    // sub-tables of at least 1024 slots, at most 256 of them (single-word keys only); more
    // of them when that lets a sub-table fit the LDS, so that the two-phase path folds
    // straight into its sub-tables.  Cold rows go to at most kMaxRecParts partitions, each
    // holding the rows of up to 2^kMaxSplitBits sub-tables.
    int32_t parts = 1;
    if (h->ks.kw == 1) {
        while (parts < 256 && cap / (uint64_t)(parts * 2) >= 1024) parts *= 2;
        const size_t budget = std::min<size_t>(device_info().lds_per_block, 160 * 1024);
        int32_t more = parts;
        while (more < (kMaxRecParts << kMaxSplitBits) &&
               p2o_lds_bytes((int64_t)(cap / (uint64_t)more), h->fields) > budget &&
               cap / (uint64_t)(more * 2) >= 1024)
            more *= 2;
        if (p2o_lds_bytes((int64_t)(cap / (uint64_t)more), h->fields) <= budget) parts = more;
    }
    h->parts = parts;
    h->rec_parts = std::min<int32_t>(parts, kMaxRecParts);
}

// Slot arrays for h->cap (+ 2 side slots); on failure every array is released.
int alloc_slots(qk_groupby_s* h) {
    const size_t slots = (size_t)(h->cap + 2);
    const uint32_t fields = h->fields;
    int st = QK_SUCCESS;
    auto A = [&](unsigned long long** p, size_t bytes) {
        if (st == QK_SUCCESS) st = pool_alloc((void**)p, bytes);
    };
    h->keys = nullptr;
    h->cnt = h->cntall = h->sum = h->mn = h->mx = nullptr;
    A((unsigned long long**)&h->keys, slots * h->ks.kw * 8);
    A(&h->cnt, slots * 8);
    if (fields & F_CNTALL) A(&h->cntall, slots * 8);
    if (fields & F_SUM) A(&h->sum, slots * 8);
    if (fields & F_MIN) A(&h->mn, slots * 8);
    if (fields & F_MAX) A(&h->mx, slots * 8);
    if (st != QK_SUCCESS) {
        void* ptrs[] = {h->keys, h->cnt, h->cntall, h->sum, h->mn, h->mx};
        for (void* q : ptrs)
            if (q) pool_free(q);
        h->keys = nullptr;
        h->cnt = h->cntall = h->sum = h->mn = h->mx = nullptr;
    }
    return st;
}

bool agg_needs_value(int a) {
    return a == QK_AGG_MIN || a == QK_AGG_MAX || a == QK_AGG_SUM || a == QK_AGG_AVG;
}

int init_table(qk_groupby_s* h, hipStream_t s) {
    GView g = view_of(h);
    const uint64_t total = h->cap + 2;
    int64_t grid = (int64_t)((total + 255) / 256);
    const int64_t capg = (int64_t)device_info().cus * 16;
    if (grid > capg) grid = capg;
    k_gb_init<<<(unsigned)grid, 256, 0, s>>>(g, h->fields);
    QK_HIP(hipGetLastError());
    if (h->seed_slot || h->seed_bits) {  // a reset drops the seeded keys too
        QK_HIP(hipStreamSynchronize(s));
        if (h->seed_slot) pool_free(h->seed_slot);
        if (h->seed_bits) pool_free(h->seed_bits);
        h->seed_slot = nullptr;
        h->seed_bits = nullptr;
        h->seed_n = 0;
    }
    h->tp.heavy_ready = false;
    h->tp.pending_rows = 0;  // rows of the previous query still in the regions are dropped
    return QK_SUCCESS;
}

// Merge rows whose phase 2 was deferred; everything that reads the table calls this first.
int flush_pending(qk_groupby_s* h, hipStream_t s) {
    if (h->tp.pending_rows == 0) return QK_SUCCESS;
    return twophase_flush(h->tp, view_of(h), h->fields, s);
}

int read_counters(qk_groupby_s* h, hipStream_t s, TableCounters* out) {
    void* hb = nullptr;
    QK_TRY(pinned_host(sizeof(TableCounters), &hb));
    QK_HIP(hipMemcpyAsync(hb, h->ctr, sizeof(TableCounters), hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    memcpy(out, hb, sizeof(TableCounters));
    return QK_SUCCESS;
}

// Read and clear the latched error (ErrorFlagSwapper::Swap, ErrorFlagSwapper.h:32-48).
int take_error(qk_groupby_s* h, hipStream_t s) {
    TableCounters c;
    QK_TRY(read_counters(h, s, &c));
    if (c.err != 0) {
        QK_HIP(hipMemsetAsync(&h->ctr->err, 0, sizeof(int32_t), s));
        if (c.err == QK_HASH_TABLE_FULL)
            return fail(QK_HASH_TABLE_FULL, "GROUP BY hash table is full (%llu keys, capacity %llu)",
                        (unsigned long long)c.nkeys, (unsigned long long)h->logical_cap);
        return fail(c.err, "GROUP BY kernel error %d", c.err);
    }
    return QK_SUCCESS;
}

// Grow the table until it holds at least `need` keys (at least doubling it): new slot arrays,
// every group moved over by the merge kernel (a key keeps its states), the old arrays freed.
// Deferred rows are merged into the old table first.  Waits for the table's work.
int grow_table(qk_groupby_s* h, uint64_t need, hipStream_t s) {
    if (h->logical_cap >= kMaxLogical)
        return fail(QK_HASH_TABLE_FULL, "GROUP BY hash table is full at its largest size (%llu keys)",
                    (unsigned long long)h->logical_cap);
    QK_TRY(flush_pending(h, s));
    const GView old = view_of(h);
    qk_groupby_s keep = {};
    keep.keys = h->keys;
    keep.cnt = h->cnt;
    keep.cntall = h->cntall;
    keep.sum = h->sum;
    keep.mn = h->mn;
    keep.mx = h->mx;
    const uint64_t old_logical = h->logical_cap;
    const uint64_t old_cap = h->cap;
    const int32_t old_parts = h->parts, old_rec = h->rec_parts, old_max = h->reported_max;
    uint64_t logical = old_logical * 2;
    while (logical < need && logical < kMaxLogical) logical *= 2;
    layout_table(h, std::min<uint64_t>(logical, kMaxLogical));
    int st = alloc_slots(h);
    if (st != QK_SUCCESS) {  // the table stays as it was
        h->keys = keep.keys;
        h->cnt = keep.cnt;
        h->cntall = keep.cntall;
        h->sum = keep.sum;
        h->mn = keep.mn;
        h->mx = keep.mx;
        h->logical_cap = old_logical;
        h->cap = old_cap;
        h->parts = old_parts;
        h->rec_parts = old_rec;
        h->reported_max = old_max;
        return st;
    }
    GView g = view_of(h);
    const uint64_t total = h->cap + 2;
    const unsigned igrid = (unsigned)std::min<uint64_t>((total + 255) / 256, (uint64_t)device_info().cus * 16);
    k_gb_init<<<igrid, 256, 0, s>>>(g, h->fields);
    QK_HIP(hipGetLastError());
    const unsigned mgrid = (unsigned)std::min<uint64_t>((old.cap + 2 + 255) / 256, (uint64_t)device_info().cus * 16);
    switch (h->ks.kw) {
    case 1: k_gb_merge<1><<<mgrid, 256, 0, s>>>(g, old, h->fields); break;
    case 2: k_gb_merge<2><<<mgrid, 256, 0, s>>>(g, old, h->fields); break;
    default: k_gb_merge<4><<<mgrid, 256, 0, s>>>(g, old, h->fields); break;
    }
    QK_HIP(hipGetLastError());
    QK_HIP(hipStreamSynchronize(s));
    void* ptrs[] = {keep.keys, keep.cnt, keep.cntall, keep.sum, keep.mn, keep.mx};
    for (void* q : ptrs)
        if (q) pool_free(q);
    h->grows++;
    // the two-phase buffers are laid out for the partition count: the next block resizes them
    // (and samples its heavy keys again) when that changed
    return QK_SUCCESS;
}

int choose_lds_slots(const qk_groupby_s* h) {
    if (h->ks.kw != 1 || h->lds_budget == 0) return 0;
    const DeviceInfo& d = device_info();
    size_t budget = h->lds_budget > 0 ? (size_t)h->lds_budget : d.lds_per_cu / 2;
    if (budget > d.lds_per_block) budget = d.lds_per_block;
    const int bps = lds_slot_bytes(h->fields);
    int64_t slots = (int64_t)(budget / (size_t)bps) - 2;
    // no point holding more LDS slots than twice the table's key capacity
    const int64_t useful = (int64_t)std::min<uint64_t>(h->logical_cap * 2 + 64, (uint64_t)1 << 30);
    if (slots > useful) slots = useful;
    if (slots < 64) return 0;
    return (int)slots;
}

std::pair<hipEvent_t, hipEvent_t>* next_events(qk_groupby h) {
    if (h->events_used == h->events.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
        h->events.emplace_back(a, b);
    }
    return &h->events[h->events_used++];
}

int process_once(qk_groupby h, int mode, const void* const* keys, const uint64_t* const* key_nulls,
                 const void* values, const uint64_t* value_nulls, const int8_t* mask,
                 int32_t npreds, const qk_predicate* preds, int32_t n, void* stream) {
    if (!h) return fail(QK_ERROR, "group by: null table");
    QK_TRY(ensure_init());
    if (n <= 0) return QK_SUCCESS;  // GPUGroupBy.cuh:401
    if (!keys) return fail(QK_ERROR, "group by: null key array");
    KeyCols kc{};
    for (int c = 0; c < h->nkeys; c++) {
        if (!keys[c]) return fail(QK_ERROR, "group by: null key column %d", c);
        kc.col[c] = keys[c];
        kc.nm[c] = key_nulls ? key_nulls[c] : nullptr;
        if (kc.nm[c] && (h->flags & QK_GB_KEYS_NOT_NULL))
            return fail(QK_ERROR, "group by: key null mask given to a table created with QK_GB_KEYS_NOT_NULL");
    }
    bool needs_value = false;
    for (int a = 0; a < h->naggs; a++) needs_value = needs_value || agg_needs_value(h->aggs[a]);
    if (needs_value && !values) return fail(QK_ERROR, "group by: aggregate needs a value column");
    if (mode == 1 && !mask) return fail(QK_ERROR, "group by: null mask");
    PredSet ps{};
    QK_TRY(check_preds(npreds, preds, &ps));
    uint32_t vpred = 0;
    for (int p = 0; p < ps.n; p++)
        if (values && ps.p[p].column == values && ps.p[p].type == h->value_type &&
            ps.p[p].null_mask == value_nulls)
            vpred |= 1u << p;
    // WHERE on the value column alone: test one interval per row instead of the list
    if (mode == 2 && values && fold_value_range(&ps, vpred, h->value_type)) mode = 3;
    else if (mode == 2 && fold_column_range(&ps, vpred)) mode = 4;
    hipStream_t s = as_stream(stream);
    GView g = view_of(h);
    std::pair<hipEvent_t, hipEvent_t>* ev = h->profiling ? next_events(h) : nullptr;
    const int vtype = values ? h->value_type : 0;
    const bool two = twophase_supported(h->ks.kw, h->value_type) &&
                     (h->twophase_mode == 1 || (h->twophase_mode < 0 && n >= h->twophase_min_rows));
    int st;
    if (two) {
        // Phase 2 is deferred across blocks only for asynchronous tables: a synchronous table
        // reports a full table at the block that filled it, as the reference does, and a
        // growing table grows at the block that filled it.
        const bool defer = (h->flags & (QK_GB_ASYNC | QK_GB_AUTO_GROW)) == QK_GB_ASYNC && h->tp.defer_rows != 0;
        TwoPhaseBufs& tp = h->tp;
        if (tp.pending_rows > 0 &&
            (twophase_needs_realloc(tp, h->rec_parts, h->value_type, values != nullptr, n, defer) ||
             tp.pending_rows + n > tp.rows_alloc || tp.pend_vtype != vtype))
            QK_TRY(twophase_flush(tp, g, h->fields, s));
        QK_TRY(twophase_prepare(tp, h->rec_parts, h->fields, h->value_type, values != nullptr, n, defer, s));
        if (!tp.heavy_ready)
            QK_TRY(twophase_sample(tp, g, h->ks, kc, vtype, values, value_nulls, mask, ps, vpred, n, mode, s));
        h->last_H = tp.H1;
        h->last_path = 1;
        st = twophase_phase1(tp, g, h->ks, kc, vtype, values, value_nulls, mask, ps, vpred, n,
                             h->fields, mode, s, ev ? ev->first : nullptr, ev ? ev->second : nullptr);
        if (st == QK_SUCCESS && !defer) st = twophase_flush(tp, g, h->fields, s);
    } else {
        const int32_t H = choose_lds_slots(h);
        h->last_H = H;
        h->last_path = 0;
        if (ev) QK_HIP(hipEventRecord(ev->first, s));
        switch (h->ks.kw) {
        case 1: st = launch_rows_v<1>(vtype, mode, g, h->ks, kc, values, value_nulls, mask, ps, vpred, n, h->fields, H, h->block, h->wgs_per_cu, s); break;
        case 2: st = launch_rows_v<2>(vtype, mode, g, h->ks, kc, values, value_nulls, mask, ps, vpred, n, h->fields, 0, h->block, h->wgs_per_cu, s); break;
        default: st = launch_rows_v<4>(vtype, mode, g, h->ks, kc, values, value_nulls, mask, ps, vpred, n, h->fields, 0, h->block, h->wgs_per_cu, s); break;
        }
        if (st == QK_SUCCESS && ev) QK_HIP(hipEventRecord(ev->second, s));
    }
    if (st != QK_SUCCESS) return st;
    if (!(h->flags & (QK_GB_ASYNC | QK_GB_AUTO_GROW))) return take_error(h, s);
    return QK_SUCCESS;
}

// Rows of the block not yet aggregated by a table that dropped partials: an int8 mask of the
// rows that pass the block's WHERE and whose key the table lacks.
int absent_rows(qk_groupby h, int mode, const KeyCols& kc, const int8_t* mask, const PredSet& ps,
                int32_t n, int8_t* out, hipStream_t s) {
    const GView g = view_of(h);
    const unsigned grid = (unsigned)std::min<int64_t>(((int64_t)n + 255) / 256, (int64_t)device_info().cus * 16);
#define QK_ABSENT(KW, M) k_gb_absent<KW, M><<<grid, 256, 0, s>>>(g, h->ks, kc, mask, ps, n, out)
    switch (h->ks.kw * 4 + mode) {
    case 4: QK_ABSENT(1, 0); break;
    case 5: QK_ABSENT(1, 1); break;
    case 6: QK_ABSENT(1, 2); break;
    case 8: QK_ABSENT(2, 0); break;
    case 9: QK_ABSENT(2, 1); break;
    case 10: QK_ABSENT(2, 2); break;
    case 16: QK_ABSENT(4, 0); break;
    case 17: QK_ABSENT(4, 1); break;
    default: QK_ABSENT(4, 2); break;
    }
#undef QK_ABSENT
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

// ProcessBlock.  A QK_GB_AUTO_GROW table checks itself after the block: past its key capacity it
// grows (GpuSqlDispatcher.cpp:378-394 grows and restarts the query instead), and when the block
// found sub-tables full, the rows whose keys did not get in are aggregated again into the grown
// table.  Those are exactly the rows whose key the table lacks (g_set_lost), so no row is
// counted twice and none is missed.
int process_common(qk_groupby h, int mode, const void* const* keys, const uint64_t* const* key_nulls,
                   const void* values, const uint64_t* value_nulls, const int8_t* mask,
                   int32_t npreds, const qk_predicate* preds, int32_t n, void* stream) {
    int st = process_once(h, mode, keys, key_nulls, values, value_nulls, mask, npreds, preds, n, stream);
    if (st != QK_SUCCESS || !(h->flags & QK_GB_AUTO_GROW) || n <= 0) return st;
    hipStream_t s = as_stream(stream);
    KeyCols kc{};
    for (int c = 0; c < h->nkeys; c++) {
        kc.col[c] = keys[c];
        kc.nm[c] = key_nulls ? key_nulls[c] : nullptr;
    }
    PredSet ps{};
    QK_TRY(check_preds(npreds, preds, &ps));
    // rows to aggregate again, rewritten in place each round (a row's entry only ever clears)
    int8_t* again = nullptr;
    for (;;) {
        TableCounters c;
        st = read_counters(h, s, &c);
        if (st != QK_SUCCESS || c.err == 0) break;
        QK_HIP(hipMemsetAsync(&h->ctr->err, 0, 2 * sizeof(int32_t), s));  // err and lost
        if (c.err != QK_HASH_TABLE_FULL) {
            st = fail(c.err, "GROUP BY kernel error %d", c.err);
            break;
        }
        st = grow_table(h, c.nkeys + 1, s);
        if (st != QK_SUCCESS) break;
        if (!c.lost) continue;
        if (!again) {
            st = pool_alloc((void**)&again, (size_t)n);
            if (st != QK_SUCCESS) break;
        }
        st = absent_rows(h, mode, kc, mask, ps, n, again, s);
        if (st == QK_SUCCESS)
            st = process_once(h, 1, keys, key_nulls, values, value_nulls, again, 0, nullptr, n, stream);
        if (st != QK_SUCCESS) break;
        mode = 1;
        mask = again;
    }
    if (again) {
        (void)hipStreamSynchronize(s);
        pool_free(again);
    }
    return st;
}

int collect_profile(qk_groupby h) {
    for (size_t i = 0; i < h->events_used; i++) {
        QK_HIP(hipEventSynchronize(h->events[i].second));
        float ms = 0.f;
        QK_HIP(hipEventElapsedTime(&ms, h->events[i].first, h->events[i].second));
        h->prof_ms += ms;
        h->prof_launches++;
    }
    h->events_used = 0;
    return QK_SUCCESS;
}

void free_table(qk_groupby_s* h) {
    void* ptrs[] = {h->keys, h->cnt, h->cntall, h->sum, h->mn, h->mx, h->ctr, h->seed_slot, h->seed_bits};
    for (void* p : ptrs)
        if (p) pool_free(p);
    twophase_free(h->tp);
}

}  // namespace
}  // namespace qk

extern "C" {

int qk_gb_create(qk_groupby* out, int32_t nkeys, const int32_t* key_types, int32_t value_type,
                 int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                 int32_t max_hash_count, int32_t key_size_multiplier, uint32_t flags) {
    if (!out) return fail(QK_ERROR, "qk_gb_create: null output");
    *out = nullptr;
    QK_TRY(ensure_init());
    if (nkeys < 1 || nkeys > kMaxKeys) return fail(QK_ERROR, "qk_gb_create: 1..%d key columns", kMaxKeys);
    if (naggs < 1 || naggs > kMaxAggs) return fail(QK_ERROR, "qk_gb_create: 1..%d aggregates", kMaxAggs);
    if (!key_types || !aggs) return fail(QK_ERROR, "qk_gb_create: null type/aggregate array");
    if (max_hash_count < 1 || key_size_multiplier < 1)
        return fail(QK_ERROR, "qk_gb_create: capacity must be positive");
    std::unique_ptr<qk_groupby_s> h(new qk_groupby_s());
    h->nkeys = nkeys;
    h->value_type = value_type;
    h->naggs = naggs;
    h->flags = flags;
    bool needs_value = false;
    uint32_t fields = F_CNT;
    for (int a = 0; a < naggs; a++) {
        const int ag = aggs[a];
        if (ag < QK_AGG_MIN || ag > QK_AGG_COUNT_ALL)
            return fail(QK_UNKNOWN_AGG_FUN, "qk_gb_create: unknown aggregate %d", ag);
        h->aggs[a] = ag;
        if (agg_needs_value(ag)) needs_value = true;
        if (ag == QK_AGG_SUM || ag == QK_AGG_AVG) fields |= F_SUM;
        if (ag == QK_AGG_MIN) fields |= F_MIN;
        if (ag == QK_AGG_MAX) fields |= F_MAX;
        if (ag == QK_AGG_COUNT_ALL) fields |= F_CNTALL;
    }
    if (needs_value && value_type != QK_INT && value_type != QK_LONG && value_type != QK_FLOAT &&
        value_type != QK_DOUBLE)
        return fail(QK_ERROR, "qk_gb_create: value type %d unsupported", value_type);
    if (!needs_value && value_type != QK_INT && value_type != QK_LONG && value_type != QK_FLOAT &&
        value_type != QK_DOUBLE)
        value_type = h->value_type = QK_INT;  // no value column is read; any placeholder type
    for (int a = 0; a < naggs; a++) {
        const int ag = aggs[a];
        int dflt = (ag == QK_AGG_COUNT || ag == QK_AGG_COUNT_ALL) ? QK_LONG : value_type;
        int ot = (out_types && out_types[a] > 0) ? out_types[a] : dflt;
        if (ag == QK_AGG_COUNT || ag == QK_AGG_COUNT_ALL) {
            if (ot != QK_LONG)  // GPUGroupBy.cuh:563-567
                return fail(QK_ERROR, "Output value data type in GROUP BY with COUNT must be int64_t");
        } else if (ag == QK_AGG_SUM) {
            bool widen = (value_type == QK_INT && ot == QK_LONG) || (value_type == QK_FLOAT && ot == QK_DOUBLE);
            if (ot != value_type && !widen)
                return fail(QK_ERROR, "Input and output value data type must be the same in GROUP BY");
        } else if (ag != QK_AGG_NONE && ot != value_type) {  // GPUGroupBy.cuh:524-529
            return fail(QK_ERROR, "Input and output value data type must be the same in GROUP BY");
        }
        h->out_types[a] = ot;
    }
    h->fields = fields;
    h->real_sum = (value_type == QK_FLOAT || value_type == QK_DOUBLE);

    // key layout
    KeySpec ks{};
    ks.nkeys = nkeys;
    for (int c = 0; c < nkeys; c++) {
        const int t = key_types[c];
        if (!type_size(t)) return fail(QK_ERROR, "qk_gb_create: key type %d unsupported", t);
        h->key_types[c] = t;
        ks.type[c] = t;
    }
    if (nkeys == 1) {
        ks.kw = 1;
        ks.single = 1;
    } else {
        int pos = 0;
        for (int c = 0; c < nkeys; c++) {
            const int w = (key_types[c] == QK_LONG || key_types[c] == QK_DOUBLE) ? 64 : 32;
            if ((pos % 64) + w > 64) pos = (pos / 64 + 1) * 64;
            ks.word[c] = pos / 64;
            ks.shift[c] = pos % 64;
            pos += w;
        }
        if ((flags & QK_GB_KEYS_NOT_NULL) && pos <= 64) {
            ks.kw = 1;
        } else {
            ks.nullable = (flags & QK_GB_KEYS_NOT_NULL) ? 0 : 1;
            if (ks.nullable)
                for (int c = 0; c < nkeys; c++) {
                    ks.nword[c] = pos / 64;
                    ks.nshift[c] = pos % 64;
                    pos++;
                }
            const int kw = (pos + 1 + 63) / 64;  // + the present bit (bit 63 of the last word)
            if (pos > kw * 64 - 1) return fail(QK_ERROR, "qk_gb_create: key too wide");
            if (kw > 4) return fail(QK_ERROR, "qk_gb_create: composite key wider than 255 bits");
            ks.kw = kw == 3 ? 4 : kw;
        }
    }
    h->ks = ks;

    layout_table(h.get(), std::min<uint64_t>((uint64_t)max_hash_count * (uint64_t)key_size_multiplier, kMaxLogical));
    int st = alloc_slots(h.get());
    if (st == QK_SUCCESS) st = pool_alloc((void**)&h->ctr, sizeof(TableCounters));
    if (st == QK_SUCCESS) {
        hipError_t e = hipGetDevice(&h->device);
        if (e != hipSuccess) st = fail(QK_RUNTIME_ERROR, "hipGetDevice: %s", hipGetErrorString(e));
    }
    if (st == QK_SUCCESS) st = init_table(h.get(), nullptr);
    if (st == QK_SUCCESS) {
        hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) st = fail(QK_RUNTIME_ERROR, "table init: %s", hipGetErrorString(e));
    }
    if (st != QK_SUCCESS) {
        free_table(h.get());
        return st;
    }
    *out = h.release();
    return QK_SUCCESS;
}

int qk_gb_destroy(qk_groupby h) {
    if (!h) return QK_SUCCESS;
    (void)hipDeviceSynchronize();
    free_table(h);
    for (auto& e : h->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    delete h;
    return QK_SUCCESS;
}

int qk_gb_reset(qk_groupby h, void* stream) {
    if (!h) return fail(QK_ERROR, "qk_gb_reset: null table");
    return init_table(h, as_stream(stream));
}

int qk_gb_get_max_hash_count(qk_groupby h, int32_t* out) {
    if (!h || !out) return fail(QK_ERROR, "qk_gb_get_max_hash_count: null argument");
    *out = h->reported_max;
    return QK_SUCCESS;
}

int qk_gb_process_block(qk_groupby h, const void* const* keys, const uint64_t* const* key_nulls,
                        const void* values, const uint64_t* value_nulls, int32_t n, void* stream) {
    return process_common(h, 0, keys, key_nulls, values, value_nulls, nullptr, 0, nullptr, n, stream);
}

int qk_gb_process_block_masked(qk_groupby h, const void* const* keys,
                               const uint64_t* const* key_nulls, const void* values,
                               const uint64_t* value_nulls, const int8_t* mask, int32_t n,
                               void* stream) {
    return process_common(h, 1, keys, key_nulls, values, value_nulls, mask, 0, nullptr, n, stream);
}

int qk_gb_process_block_where(qk_groupby h, const void* const* keys,
                              const uint64_t* const* key_nulls, const void* values,
                              const uint64_t* value_nulls, int32_t n, int32_t npreds,
                              const qk_predicate* preds, void* stream) {
    return process_common(h, npreds > 0 ? 2 : 0, keys, key_nulls, values, value_nulls, nullptr,
                          npreds, preds, n, stream);
}

int qk_gb_check_error(qk_groupby h, void* stream) {
    if (!h) return fail(QK_ERROR, "qk_gb_check_error: null table");
    QK_TRY(flush_pending(h, as_stream(stream)));
    return take_error(h, as_stream(stream));
}

int qk_gb_group_count(qk_groupby h, int64_t* out, void* stream) {
    if (!h || !out) return fail(QK_ERROR, "qk_gb_group_count: null argument");
    hipStream_t s = as_stream(stream);
    QK_TRY(flush_pending(h, s));
    TableCounters c;
    QK_TRY(read_counters(h, s, &c));
    uint64_t side = 0;
    if (h->ks.kw == 1) {
        void* hb = nullptr;
        QK_TRY(pinned_host(8, &hb));
        QK_HIP(hipMemcpyAsync(hb, h->keys + h->cap, 8, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        uint64_t w;
        memcpy(&w, hb, 8);
        side = w != 0 ? 1 : 0;  // the NULL group
    }
    *out = (int64_t)(c.nkeys + side);
    return QK_SUCCESS;
}

}  // extern "C"

namespace qk {
namespace {

// GetResults with the outputs `aggs` (naggs of them, output types otypes) computed from the
// table's aggregate states.
int get_results_impl(qk_groupby h, int32_t naggs, const int32_t* aggs, const int32_t* otypes,
                     void** out_keys, uint64_t** out_key_nulls, void** out_values,
                     uint64_t** out_value_nulls, int32_t* out_count, void* stream) {
    if (!h || !out_count) return fail(QK_ERROR, "qk_gb_get_results: null argument");
    if (!out_keys) return fail(QK_ERROR, "qk_gb_get_results: null key output array");
    QK_TRY(ensure_init());
    hipStream_t s = as_stream(stream);
    *out_count = 0;
    for (int c = 0; c < h->nkeys; c++) {
        out_keys[c] = nullptr;
        if (out_key_nulls) out_key_nulls[c] = nullptr;
    }
    for (int a = 0; a < naggs; a++) {
        if (out_values) out_values[a] = nullptr;
        if (out_value_nulls) out_value_nulls[a] = nullptr;
    }
    QK_TRY(flush_pending(h, s));
    // the latched error, the key count and (single-word keys) the NULL group's slot word in
    // one round trip to the host
    TableCounters c;
    bool null_group = false;
    {
        void* hb = nullptr;
        QK_TRY(pinned_host(64, &hb));
        QK_HIP(hipMemcpyAsync(hb, h->ctr, sizeof(TableCounters), hipMemcpyDeviceToHost, s));
        if (h->ks.kw == 1)
            QK_HIP(hipMemcpyAsync((char*)hb + 32, h->keys + h->cap, 8, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        static_assert(sizeof(TableCounters) <= 32, "counters and slot word share the staging buffer");
        memcpy(&c, hb, sizeof(TableCounters));
        if (h->ks.kw == 1) {
            uint64_t w;
            memcpy(&w, (char*)hb + 32, 8);
            null_group = w != 0;
        }
    }
    if (c.err != 0) {  // read and cleared, as take_error does
        QK_HIP(hipMemsetAsync(&h->ctr->err, 0, sizeof(int32_t), s));
        if (c.err == QK_HASH_TABLE_FULL)
            return fail(QK_HASH_TABLE_FULL, "GROUP BY hash table is full (%llu keys, capacity %llu)",
                        (unsigned long long)c.nkeys, (unsigned long long)h->logical_cap);
        return fail(c.err, "GROUP BY kernel error %d", c.err);
    }
    const int64_t groups = (int64_t)c.nkeys + (null_group ? 1 : 0);
    if (groups == 0) return QK_SUCCESS;
    if (groups > INT32_MAX) return fail(QK_ERROR, "qk_gb_get_results: too many groups");

    OutSpec os{};
    os.naggs = naggs;
    os.vtype = h->value_type;
    std::vector<void*> owned;
    auto cleanup = [&]() {
        for (void* p : owned) pool_free(p);
    };
    auto alloc = [&](size_t bytes, void** p) -> int {
        int st = pool_alloc(p, bytes);
        if (st == QK_SUCCESS) owned.push_back(*p);
        return st;
    };
    const size_t nwords = (size_t)((groups + 63) / 64);
    int st = QK_SUCCESS;
    for (int c = 0; c < h->nkeys && st == QK_SUCCESS; c++) {
        st = alloc((size_t)groups * type_size(h->key_types[c]), &os.kout[c]);
        if (st == QK_SUCCESS && out_key_nulls) st = alloc(nwords * 8, (void**)&os.konull[c]);
    }
    for (int a = 0; a < naggs && st == QK_SUCCESS; a++) {
        os.agg[a] = aggs[a];
        os.otype[a] = otypes[a];
        if (aggs[a] == QK_AGG_NONE || !out_values) continue;
        st = alloc((size_t)groups * type_size(otypes[a]), &os.out[a]);
        if (st == QK_SUCCESS && out_value_nulls) st = alloc(nwords * 8, (void**)&os.onull[a]);
    }
    if (st == QK_SUCCESS) {
        // every null mask cleared by one launch (emit only sets bits)
        ZeroList zl{};
        for (int c = 0; c < h->nkeys; c++)
            if (os.konull[c]) zl.p[zl.n++] = os.konull[c];
        for (int a = 0; a < naggs; a++)
            if (os.onull[a]) zl.p[zl.n++] = os.onull[a];
        if (zl.n) {
            const int64_t grid = std::min<int64_t>(((int64_t)nwords + 255) / 256, (int64_t)device_info().cus * 4);
            k_zero_words<<<dim3((unsigned)grid, (unsigned)zl.n), 256, 0, s>>>(zl, (int64_t)nwords);
            if (hipGetLastError() != hipSuccess) st = QK_RUNTIME_ERROR;
        }
    }
    if (st != QK_SUCCESS) {
        cleanup();
        return st == QK_RUNTIME_ERROR ? fail(st, "qk_gb_get_results: clearing the null masks failed") : st;
    }
    const int64_t nslots = (int64_t)h->cap + 2;
    const int64_t tiles = (nslots + kEmitTile - 1) / kEmitTile;
    Scratch* sc = nullptr;
    st = stream_scratch(s, lb_scratch_bytes(tiles) + 64, &sc);
    if (st != QK_SUCCESS) {
        cleanup();
        return st;
    }
    uint64_t* words = (uint64_t*)sc->ptr;
    int32_t* total = (int32_t*)(words + kLbHeaderWords + tiles);
    const uint32_t epoch = ++sc->epoch;
    GView g = view_of(h);
    k_gb_emit<<<(unsigned)tiles, kEmitBlock, 0, s>>>(g, h->ks, os, words + kLbHeaderWords,
                                                     (uint32_t*)words, epoch, tiles, total,
                                                     groups, h->seed_bits, h->seed_n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && h->seed_n > 0) {
        const int64_t grid = std::min<int64_t>((h->seed_n + 255) / 256, (int64_t)device_info().cus * 8);
        k_gb_emit_seeded<<<(unsigned)grid, 256, 0, s>>>(g, h->ks, os, h->seed_slot, h->seed_n);
        e = hipGetLastError();
    }
    int32_t written = 0;
    if (e == hipSuccess) {
        void* hb = nullptr;
        st = pinned_host(8, &hb);
        if (st == QK_SUCCESS) {
            e = hipMemcpyAsync(hb, total, 4, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e == hipSuccess) written = *(int32_t*)hb;
        }
    }
    if (e != hipSuccess || st != QK_SUCCESS) {
        cleanup();
        return st != QK_SUCCESS ? st : fail(QK_RUNTIME_ERROR, "qk_gb_get_results: %s", hipGetErrorString(e));
    }
    written += h->seed_n;
    if (written != (int32_t)groups) {
        cleanup();
        return fail(QK_ERROR, "qk_gb_get_results: %d groups emitted, %lld counted", written, (long long)groups);
    }
    for (int c = 0; c < h->nkeys; c++) {
        out_keys[c] = os.kout[c];
        if (out_key_nulls) out_key_nulls[c] = os.konull[c];
    }
    for (int a = 0; a < naggs; a++) {
        if (out_values) out_values[a] = os.out[a];
        if (out_value_nulls) out_value_nulls[a] = os.onull[a];
    }
    *out_count = (int32_t)groups;
    return QK_SUCCESS;
}

}  // namespace
}  // namespace qk

extern "C" {

int qk_gb_get_results(qk_groupby h, void** out_keys, uint64_t** out_key_nulls, void** out_values,
                      uint64_t** out_value_nulls, int32_t* out_count, void* stream) {
    if (!h) return fail(QK_ERROR, "qk_gb_get_results: null table");
    return get_results_impl(h, h->naggs, h->aggs, h->out_types, out_keys, out_key_nulls, out_values,
                            out_value_nulls, out_count, stream);
}

int qk_gb_get_results_aggs(qk_groupby h, int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                           void** out_keys, uint64_t** out_key_nulls, void** out_values,
                           uint64_t** out_value_nulls, int32_t* out_count, void* stream) {
    if (!h) return fail(QK_ERROR, "qk_gb_get_results_aggs: null table");
    if (naggs < 1 || naggs > kMaxAggs || !aggs) return fail(QK_ERROR, "qk_gb_get_results_aggs: 1..%d aggregates", kMaxAggs);
    int32_t ot[kMaxAggs];
    for (int a = 0; a < naggs; a++) {
        const int ag = aggs[a];
        uint32_t need = 0;
        if (ag == QK_AGG_SUM || ag == QK_AGG_AVG) need = F_SUM;
        else if (ag == QK_AGG_MIN) need = F_MIN;
        else if (ag == QK_AGG_MAX) need = F_MAX;
        else if (ag == QK_AGG_COUNT_ALL) need = F_CNTALL;
        else if (ag != QK_AGG_COUNT && ag != QK_AGG_NONE) return fail(QK_UNKNOWN_AGG_FUN, "unknown aggregate %d", ag);
        if ((h->fields & need) != need)
            return fail(QK_ERROR, "qk_gb_get_results_aggs: aggregate %d needs a state the table does not keep", ag);
        const int dflt = (ag == QK_AGG_COUNT || ag == QK_AGG_COUNT_ALL) ? QK_LONG : h->value_type;
        ot[a] = (out_types && out_types[a] > 0) ? out_types[a] : dflt;
        if ((ag == QK_AGG_COUNT || ag == QK_AGG_COUNT_ALL) && ot[a] != QK_LONG)
            return fail(QK_ERROR, "Output value data type in GROUP BY with COUNT must be int64_t");
        if (!type_size(ot[a])) return fail(QK_ERROR, "qk_gb_get_results_aggs: output type %d unsupported", ot[a]);
    }
    return get_results_impl(h, naggs, aggs, ot, out_keys, out_key_nulls, out_values, out_value_nulls, out_count,
                            stream);
}

int qk_gb_seed_keys(qk_groupby h, const void* keys, int32_t n, int32_t first_key_is_null, void* stream) {
    if (!h) return fail(QK_ERROR, "qk_gb_seed_keys: null table");
    QK_TRY(ensure_init());
    if (h->nkeys != 1) return fail(QK_ERROR, "qk_gb_seed_keys: single-key tables only");
    if (n < 0 || (n > 0 && !keys)) return fail(QK_ERROR, "qk_gb_seed_keys: bad key buffer");
    if (h->seed_n > 0) return fail(QK_ERROR, "qk_gb_seed_keys: the table is already seeded");
    if (h->flags & QK_GB_AUTO_GROW)
        return fail(QK_ERROR, "qk_gb_seed_keys: a QK_GB_AUTO_GROW table moves its keys when it grows; "
                    "seeded keys need a fixed-size table");
    if (n == 0) return QK_SUCCESS;
    hipStream_t s = as_stream(stream);
    QK_TRY(flush_pending(h, s));
    TableCounters before;
    QK_TRY(read_counters(h, s, &before));
    if (before.nkeys != 0) return fail(QK_ERROR, "qk_gb_seed_keys: the table already holds groups");
    const size_t nbits = (size_t)((h->cap + 2 + 63) / 64);
    QK_TRY(pool_alloc((void**)&h->seed_slot, (size_t)n * 8));
    int st = pool_alloc((void**)&h->seed_bits, nbits * 8);
    if (st != QK_SUCCESS) {
        pool_free(h->seed_slot);
        h->seed_slot = nullptr;
        return st;
    }
    QK_HIP(hipMemsetAsync(h->seed_bits, 0, nbits * 8, s));
    const int64_t grid = std::min<int64_t>((n + 255) / 256, (int64_t)device_info().cus * 8);
    k_gb_seed<<<(unsigned)grid, 256, 0, s>>>(view_of(h), h->key_types[0], keys, n, first_key_is_null ? 1 : 0,
                                             h->seed_slot, h->seed_bits);
    QK_HIP(hipGetLastError());
    h->seed_n = n;
    st = take_error(h, s);
    TableCounters after;
    if (st == QK_SUCCESS) st = read_counters(h, s, &after);
    const uint64_t want = (uint64_t)n - (first_key_is_null ? 1 : 0);
    if (st == QK_SUCCESS && after.nkeys != want)
        st = fail(QK_ERROR, "qk_gb_seed_keys: %llu distinct keys among %llu (the seed keys must be distinct)",
                  (unsigned long long)after.nkeys, (unsigned long long)want);
    if (st != QK_SUCCESS) init_table(h, s);  // back to an empty table
    return st;
}

int qk_gb_merge(qk_groupby dst, qk_groupby src, void* stream) {
    if (!dst || !src) return fail(QK_ERROR, "qk_gb_merge: null table");
    if (dst == src) return fail(QK_ERROR, "qk_gb_merge: a table cannot be merged into itself");
    QK_TRY(ensure_init());
    bool same = dst->nkeys == src->nkeys && dst->value_type == src->value_type && dst->fields == src->fields &&
                dst->ks.kw == src->ks.kw && dst->ks.nullable == src->ks.nullable && dst->real_sum == src->real_sum;
    for (int c = 0; same && c < dst->nkeys; c++) same = dst->key_types[c] == src->key_types[c];
    if (!same) return fail(QK_ERROR, "qk_gb_merge: the tables differ in key or aggregate layout");
    int cur = -1;
    QK_HIP(hipGetDevice(&cur));
    if (dst->device != cur || src->device != cur)
        return fail(QK_ERROR, "qk_gb_merge: both tables must live on the calling thread's device (device %d "
                    "and %d, current %d); a merge across GPUs is not built", dst->device, src->device, cur);
    hipStream_t s = as_stream(stream);
    QK_TRY(flush_pending(src, s));
    QK_TRY(flush_pending(dst, s));
    QK_TRY(take_error(src, s));
    if (dst->flags & QK_GB_AUTO_GROW) {
        // room for every group of both tables before the merge starts
        TableCounters cd, cs;
        QK_TRY(read_counters(dst, s, &cd));
        QK_TRY(read_counters(src, s, &cs));
        if (cd.nkeys + cs.nkeys > dst->logical_cap) QK_TRY(grow_table(dst, cd.nkeys + cs.nkeys, s));
    }
    GView gd = view_of(dst), gs = view_of(src);
    const int64_t grid = std::min<int64_t>((int64_t)((src->cap + 2 + 255) / 256), (int64_t)device_info().cus * 16);
    switch (dst->ks.kw) {
    case 1: k_gb_merge<1><<<(unsigned)grid, 256, 0, s>>>(gd, gs, dst->fields); break;
    case 2: k_gb_merge<2><<<(unsigned)grid, 256, 0, s>>>(gd, gs, dst->fields); break;
    default: k_gb_merge<4><<<(unsigned)grid, 256, 0, s>>>(gd, gs, dst->fields); break;
    }
    QK_HIP(hipGetLastError());
    return take_error(dst, s);
}

int qk_gb_set_profiling(qk_groupby h, int32_t enable) {
    if (!h) return fail(QK_ERROR, "qk_gb_set_profiling: null table");
    h->profiling = enable != 0;
    return QK_SUCCESS;
}

int qk_gb_get_profile(qk_groupby h, double* ms, int64_t* launches) {
    if (!h) return fail(QK_ERROR, "qk_gb_get_profile: null table");
    QK_TRY(collect_profile(h));
    if (ms) *ms = h->prof_ms;
    if (launches) *launches = h->prof_launches;
    h->prof_ms = 0.0;
    h->prof_launches = 0;
    return QK_SUCCESS;
}

int qk_gb_describe(qk_groupby h, int32_t* slot_bytes, int64_t* slots, int32_t* lds_slots) {
    if (!h) return fail(QK_ERROR, "qk_gb_describe: null table");
    if (slot_bytes) {
        int b = h->ks.kw * 8 + 8;
        if (h->cntall) b += 8;
        if (h->sum) b += 8;
        if (h->mn) b += 8;
        if (h->mx) b += 8;
        *slot_bytes = b;
    }
    if (slots) *slots = (int64_t)h->cap;
    if (lds_slots) *lds_slots = h->last_H;
    return QK_SUCCESS;
}

int qk_gb_set_tuning(qk_groupby h, int32_t block, int32_t lds_bytes, int32_t wgs_per_cu) {
    if (!h) return fail(QK_ERROR, "qk_gb_set_tuning: null table");
    if (block > 0) {
        if (block % 64 != 0 || block > 1024) return fail(QK_ERROR, "qk_gb_set_tuning: block must be a multiple of 64 up to 1024");
        h->block = block;
    }
    if (lds_bytes >= 0) h->lds_budget = lds_bytes;
    if (wgs_per_cu >= 0) h->wgs_per_cu = wgs_per_cu;
    return QK_SUCCESS;
}

int qk_gb_stats(qk_groupby h, int64_t* out, int32_t n, void* stream) {
    if (!h || !out || n < 0) return fail(QK_ERROR, "qk_gb_stats: bad argument");
    hipStream_t s = as_stream(stream);
    QK_HIP(hipStreamSynchronize(s));
    int64_t v[13] = {h->last_path, 0, 0, 0, h->tp.W, h->tp.P, h->tp.cap, h->tp.H1, h->tp.H2, h->tp.kmax, h->parts,
                     h->grows, h->tp.last_form};
    if (h->tp.heavy_meta) {
        int32_t meta[2];
        QK_HIP(hipMemcpy(meta, h->tp.heavy_meta, 8, hipMemcpyDeviceToHost));
        v[1] = meta[0];
        v[2] = meta[1];
    }
    if (h->tp.r_count) {
        std::vector<uint32_t> rc((size_t)h->tp.W * h->tp.P);
        QK_HIP(hipMemcpy(rc.data(), h->tp.r_count, rc.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t c : rc) v[3] += c;
    }
    for (int i = 0; i < n && i < 13; i++) out[i] = v[i];
    return QK_SUCCESS;
}

int qk_gb_set_path(qk_groupby h, int32_t path, int64_t min_rows) {
    if (!h) return fail(QK_ERROR, "qk_gb_set_path: null table");
    if (path < -1 || path > 1) return fail(QK_ERROR, "qk_gb_set_path: path must be -1, 0 or 1");
    h->twophase_mode = path;
    if (min_rows >= 0) h->twophase_min_rows = min_rows;
    return QK_SUCCESS;
}

int qk_gb_set_defer_rows(qk_groupby h, int64_t rows) {
    if (!h) return fail(QK_ERROR, "qk_gb_set_defer_rows: null table");
    if (rows < -1) return fail(QK_ERROR, "qk_gb_set_defer_rows: rows must be -1, 0 or positive");
    h->tp.defer_rows = rows;
    return QK_SUCCESS;
}

}  // extern "C"
