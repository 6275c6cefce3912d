// strings.hip - GROUP BY over one string key column (qk_gbs_* in include/qikkdb_amd.h), the
// counterpart of GPUGroupBy<AGG,O,std::string,V> (QueryEngine/GPUCore/GPUGroupByString.cuh:39-785).
//
// A string column is the reference's GPUString (GPUMemory.cuh:70-77): all bytes concatenated,
// and indices[i] = the end offset of string i (string i is chars[indices[i-1] .. indices[i]),
// indices[-1] = 0).
//
// Design: the strings of a table are dictionary-encoded on the device, then grouped by the
// existing numeric GROUP BY.  The dictionary is an open-addressing table of D slots; a distinct
// string owns one slot for the table's lifetime, and the slot's index is the string's int64 key
// in an inner qk_groupby, which keeps every aggregate, NULL rule and path of the numeric table.
// Per block:
//   1. k_str_probe hashes each row's bytes (FNV-1a, 64 bits: the low bits pick the home slot,
//      the high 32 bits are a tag kept in the slot word) and probes linearly.  A slot word is
//      0 (empty), tag | 1 (a string of an earlier block, its bytes in the table's arena), or
//      tag | 2^31 | r (row r of this block, claimed by CAS in this kernel).  Equal tags are
//      confirmed byte by byte against the arena or against the block's own input, which this
//      kernel never writes, so no row ever reads bytes another workgroup is still writing.
//   2. k_str_commit copies the bytes of the slots claimed in this block into the arena and
//      turns their words into arena words.
//   3. The slot column goes to the inner table with the block's values and NULL masks.
// GetResults reads the inner table's groups (slot keys) and gathers each group's string.
#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "pred.hpp"

namespace {

constexpr uint32_t kRowBit = 0x80000000u;

struct StrDict {
    unsigned long long* slots;  // [D]
    int64_t* off;               // [D] arena offset of a committed slot's string
    int32_t* len;               // [D] its length
    char* arena;
    uint64_t mask;              // D - 1
    int32_t* newlist;           // slots claimed by the current block
    unsigned int* counters;     // [0] claimed slots in the current block, [1] a row found no slot
    unsigned long long* arena_used;
};

// FNV-1a over the bytes (Fowler, Noll and Vo's 64-bit offset basis and prime).
__device__ __forceinline__ uint64_t str_hash(const char* p, int32_t len) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int32_t i = 0; i < len; i++) {
        h ^= (uint8_t)p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

__device__ __forceinline__ bool str_eq(const char* a, const char* b, int32_t len) {
    for (int32_t i = 0; i < len; i++)
        if (a[i] != b[i]) return false;
    return true;
}

__device__ __forceinline__ int64_t str_start(const int64_t* idx, int64_t i) { return i ? idx[i - 1] : 0; }

__global__ void k_str_arm(StrDict d) {
    if (threadIdx.x < 2) d.counters[threadIdx.x] = 0;
}

// Slot of every row's string (inserting new strings); -1 for a NULL key, a row the WHERE drops,
// or (with counters[1] set) a row whose string found no free slot.
__global__ void __launch_bounds__(256) k_str_probe(StrDict d, const char* __restrict__ chars,
                                                   const int64_t* __restrict__ idx,
                                                   const uint64_t* __restrict__ knull, int32_t n,
                                                   qk::PredSet preds, int64_t* __restrict__ out_slot) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int64_t res = -1;
        bool pass = !(knull && qk::null_bit(knull, i));
        for (int p = 0; p < preds.n && pass; p++) pass = qk::eval_pred(preds.p[p], i);
        if (pass) {
            const int64_t start = str_start(idx, i);
            const int32_t len = (int32_t)(idx[i] - start);
            const char* s = chars + start;
            const uint64_t h = str_hash(s, len);
            const unsigned long long tag = h & 0xffffffff00000000ull;
            uint64_t pos = h & d.mask;
            for (uint64_t pr = 0; pr <= d.mask; pr++) {
                unsigned long long w = *(volatile unsigned long long*)(d.slots + pos);
                if (w == 0ull) {
                    w = atomicCAS(d.slots + pos, 0ull, tag | kRowBit | (unsigned long long)i);
                    if (w == 0ull) {
                        d.newlist[atomicAdd(d.counters, 1u)] = (int32_t)pos;
                        res = (int64_t)pos;
                        break;
                    }
                }
                if ((w & 0xffffffff00000000ull) == tag) {
                    const uint32_t low = (uint32_t)w;
                    const char* o;
                    int32_t olen;
                    if (low & kRowBit) {
                        const int64_t r = low & ~kRowBit;
                        const int64_t os = str_start(idx, r);
                        o = chars + os;
                        olen = (int32_t)(idx[r] - os);
                    } else {
                        o = d.arena + d.off[pos];
                        olen = d.len[pos];
                    }
                    if (olen == len && str_eq(o, s, len)) {
                        res = (int64_t)pos;
                        break;
                    }
                }
                pos = (pos + 1) & d.mask;
            }
            if (res < 0) atomicOr(d.counters + 1, 1u);
        }
        out_slot[i] = res;
    }
}

// The bytes of the slots the block claimed go to the arena; their words become arena words.
__global__ void __launch_bounds__(256) k_str_commit(StrDict d, const char* __restrict__ chars,
                                                    const int64_t* __restrict__ idx) {
    const uint32_t cnt = d.counters[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
        const int32_t pos = d.newlist[k];
        const unsigned long long w = d.slots[pos];
        const int64_t r = (uint32_t)w & ~kRowBit;
        const int64_t start = str_start(idx, r);
        const int32_t len = (int32_t)(idx[r] - start);
        const int64_t off = (int64_t)atomicAdd(d.arena_used, (unsigned long long)len);
        for (int32_t b = 0; b < len; b++) d.arena[off + b] = chars[start + b];
        d.off[pos] = off;
        d.len[pos] = len;
        d.slots[pos] = (w & 0xffffffff00000000ull) | 1ull;
    }
}

// Lengths of the result strings (0 for the NULL group).
__global__ void __launch_bounds__(256) k_str_lengths(StrDict d, const int64_t* __restrict__ keys,
                                                     const uint64_t* __restrict__ knull, int32_t g,
                                                     int64_t* __restrict__ lens) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g; i += (int64_t)gridDim.x * blockDim.x) {
        const bool isnull = knull && qk::null_bit(knull, i);
        lens[i] = isnull ? 0 : d.len[keys[i]];
    }
}

// Inclusive prefix sum of int64 values in place, in three steps: per-workgroup sums of
// kScanRows values, a one-workgroup scan of those sums, then each workgroup's scan with its
// offset.
constexpr int kScanThreads = 256;
constexpr int kScanRows = kScanThreads * 4;

__device__ int64_t block_incl_scan64(int64_t v, int64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const int64_t x = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    const int64_t r = sh[t];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan64_sums(const int64_t* __restrict__ v, int64_t n,
                                                              int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanRows + (int64_t)threadIdx.x * 4;
    int64_t s = 0;
    for (int k = 0; k < 4; k++)
        if (base + k < n) s += v[base + k];
    const int64_t incl = block_incl_scan64(s, sh);
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(kScanThreads) k_scan64_top(int64_t* sums, int64_t nb) {
    __shared__ int64_t sh[kScanThreads];
    __shared__ int64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int64_t b = b0 + threadIdx.x;
        const int64_t v = b < nb ? sums[b] : 0;
        const int64_t incl = block_incl_scan64(v, sh);
        const int64_t carry = s_carry;
        if (b < nb) sums[b] = carry + incl - v;  // exclusive offset of workgroup b
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) s_carry = carry + incl;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan64_apply(int64_t* __restrict__ v, int64_t n,
                                                               const int64_t* __restrict__ sums) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanRows + (int64_t)threadIdx.x * 4;
    int64_t x[4], s = 0;
    for (int k = 0; k < 4; k++) {
        x[k] = base + k < n ? v[base + k] : 0;
        s += x[k];
    }
    int64_t run = block_incl_scan64(s, sh) - s + sums[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        run += x[k];
        if (base + k < n) v[base + k] = run;
    }
}

// Bytes of result string i from the arena.
__global__ void __launch_bounds__(256) k_str_gather(StrDict d, const int64_t* __restrict__ keys,
                                                    const uint64_t* __restrict__ knull, int32_t g,
                                                    const int64_t* __restrict__ out_idx, char* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g; i += (int64_t)gridDim.x * blockDim.x) {
        if (knull && qk::null_bit(knull, i)) continue;
        const int64_t at = str_start(out_idx, i);
        const int64_t slot = keys[i];
        const char* src = d.arena + d.off[slot];
        const int32_t len = d.len[slot];
        for (int32_t b = 0; b < len; b++) out[at + b] = src[b];
    }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

int grid_for(int64_t n, int per_block = 256, int cap = 4096) {
    const int64_t g = (n + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

}  // namespace

struct qk_gbs_s {
    qk_groupby inner = nullptr;
    uint32_t flags = 0;
    int64_t D = 0;
    unsigned long long* slots = nullptr;
    int64_t* off = nullptr;
    int32_t* len = nullptr;
    char* arena = nullptr;
    int64_t arena_cap = 0;
    int64_t arena_bound = 0;   // bytes the arena may hold (the chars of every block so far)
    int32_t* newlist = nullptr;
    int64_t newlist_cap = 0;
    int64_t* slot_col = nullptr;
    int64_t slot_cap = 0;
    unsigned int* counters = nullptr;
    unsigned long long* arena_used = nullptr;
    void* host = nullptr;      // pinned: counters and block totals read back
};

namespace {

StrDict view(qk_gbs_s* h) {
    StrDict d;
    d.slots = h->slots;
    d.off = h->off;
    d.len = h->len;
    d.arena = h->arena;
    d.mask = (uint64_t)h->D - 1;
    d.newlist = h->newlist;
    d.counters = h->counters;
    d.arena_used = h->arena_used;
    return d;
}

void gbs_free(qk_gbs_s* h) {
    for (void* p : {(void*)h->slots, (void*)h->off, (void*)h->len, (void*)h->arena, (void*)h->newlist,
                    (void*)h->slot_col, (void*)h->counters})
        if (p) qk::pool_free(p);
    if (h->host) (void)hipHostFree(h->host);
    if (h->inner) qk_gb_destroy(h->inner);
}

// Room for a block of n rows and `chars` bytes: the slot column, the claimed-slot list, and the
// arena (which grows to hold every byte of every block, so a block never overflows it).
int gbs_reserve(qk_gbs_s* h, int32_t n, int64_t chars, hipStream_t s) {
    if (h->slot_cap < n) {
        if (h->slot_col || h->newlist) QK_HIP(hipStreamSynchronize(s));
        qk::pool_free(h->slot_col);
        qk::pool_free(h->newlist);
        h->slot_col = nullptr;
        h->newlist = nullptr;
        QK_TRY(qk::pool_alloc((void**)&h->slot_col, (size_t)n * 8));
        QK_TRY(qk::pool_alloc((void**)&h->newlist, (size_t)n * 4));
        h->slot_cap = n;
        h->newlist_cap = n;
    }
    const int64_t need = h->arena_bound + chars;
    if (need > h->arena_cap) {
        const int64_t cap = std::max<int64_t>(need, std::max<int64_t>(2 * h->arena_cap, 1 << 16));
        char* a = nullptr;
        QK_TRY(qk::pool_alloc((void**)&a, (size_t)cap));
        if (h->arena) {
            QK_HIP(hipMemcpyAsync(a, h->arena, (size_t)h->arena_bound, hipMemcpyDeviceToDevice, s));
            QK_HIP(hipStreamSynchronize(s));
            qk::pool_free(h->arena);
        }
        h->arena = a;
        h->arena_cap = cap;
    }
    return QK_SUCCESS;
}

int gbs_block(qk_gbs_s* h, const char* chars, const int64_t* indices, const uint64_t* key_nulls,
              const void* values, const uint64_t* value_nulls, int32_t n, int32_t npreds,
              const qk_predicate* preds, void* stream) {
    QK_TRY(qk::ensure_init());
    if (n <= 0) return QK_SUCCESS;
    if (!indices) return qk::fail(QK_ERROR, "string group by: null indices");
    qk::PredSet ps{};
    QK_TRY(qk::check_preds(npreds, preds, &ps));
    hipStream_t s = qk::as_stream(stream);
    // bytes of the block: indices[n-1]
    int64_t* hb = (int64_t*)h->host;
    QK_HIP(hipMemcpyAsync(hb, indices + n - 1, 8, hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    const int64_t total = hb[0];
    if (total < 0) return qk::fail(QK_ERROR, "string group by: negative string offsets");
    if (total > 0 && !chars) return qk::fail(QK_ERROR, "string group by: null chars");
    QK_TRY(gbs_reserve(h, n, total, s));
    StrDict d = view(h);
    k_str_arm<<<1, 64, 0, s>>>(d);
    k_str_probe<<<grid_for(n), 256, 0, s>>>(d, chars, indices, key_nulls, n, ps, h->slot_col);
    k_str_commit<<<grid_for(n), 256, 0, s>>>(d, chars, indices);
    QK_HIP(hipGetLastError());
    h->arena_bound += total;
    unsigned int* hc = (unsigned int*)((char*)h->host + 16);
    QK_HIP(hipMemcpyAsync(hc, h->counters, 8, hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    if (hc[1]) return qk::fail(QK_HASH_TABLE_FULL, "string group by: the string dictionary is full");
    const void* kp[1] = {h->slot_col};
    const uint64_t* kn[1] = {key_nulls};
    return qk_gb_process_block_where(h->inner, kp, key_nulls ? kn : nullptr, values, value_nulls, n, npreds, preds,
                                     stream);
}

}  // namespace

extern "C" {

int qk_gbs_create(qk_gbs* out, int32_t value_type, int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                  int32_t max_hash_count, int32_t key_size_multiplier, uint32_t flags) {
    if (!out) return qk::fail(QK_ERROR, "qk_gbs_create: null output");
    *out = nullptr;
    QK_TRY(qk::ensure_init());
    if (flags & QK_GB_AUTO_GROW)
        return qk::fail(QK_ERROR, "string group by: QK_GB_AUTO_GROW is not available for string keys");
    if (max_hash_count <= 0 || key_size_multiplier <= 0)
        return qk::fail(QK_ERROR, "string group by: capacity must be positive");
    const int64_t logical = std::min<int64_t>((int64_t)max_hash_count * key_size_multiplier, (int64_t)1 << 28);
    qk_gbs_s* h = new qk_gbs_s();
    h->flags = flags;
    const int32_t kt = QK_LONG;
    int st = qk_gb_create(&h->inner, 1, &kt, value_type, naggs, aggs, out_types, (int32_t)logical, 1,
                          flags & ~(uint32_t)QK_GB_KEYS_NOT_NULL);
    int64_t D = 64;
    while (D < 2 * logical) D <<= 1;
    h->D = D;
    auto A = [&](void** p, size_t b) {
        if (st == QK_SUCCESS) st = qk::pool_alloc(p, b);
    };
    A((void**)&h->slots, (size_t)D * 8);
    A((void**)&h->off, (size_t)D * 8);
    A((void**)&h->len, (size_t)D * 4);
    A((void**)&h->counters, 64);
    if (st == QK_SUCCESS) h->arena_used = (unsigned long long*)(h->counters + 8);
    if (st == QK_SUCCESS && hipHostMalloc(&h->host, 64, hipHostMallocDefault) != hipSuccess)
        st = qk::fail(QK_OUT_OF_MEMORY, "string group by: pinned host buffer");
    if (st == QK_SUCCESS) st = qk_gbs_reset(h, nullptr);
    if (st == QK_SUCCESS) st = qk_stream_synchronize(nullptr);
    if (st != QK_SUCCESS) {
        gbs_free(h);
        delete h;
        return st;
    }
    *out = h;
    return QK_SUCCESS;
}

int qk_gbs_destroy(qk_gbs h) {
    if (!h) return QK_SUCCESS;
    (void)hipDeviceSynchronize();
    gbs_free(h);
    delete h;
    return QK_SUCCESS;
}

int qk_gbs_reset(qk_gbs h, void* stream) {
    if (!h) return qk::fail(QK_ERROR, "qk_gbs_reset: null table");
    hipStream_t s = qk::as_stream(stream);
    k_fill_u64<<<grid_for(h->D), 256, 0, s>>>(h->slots, (uint64_t)h->D);
    k_fill_u64<<<1, 64, 0, s>>>((unsigned long long*)h->counters, 8);
    QK_HIP(hipGetLastError());
    h->arena_bound = 0;
    return qk_gb_reset(h->inner, stream);
}

int qk_gbs_process_block(qk_gbs h, const char* chars, const int64_t* indices, const uint64_t* key_nulls,
                         const void* values, const uint64_t* value_nulls, int32_t n, void* stream) {
    if (!h) return qk::fail(QK_ERROR, "qk_gbs_process_block: null table");
    return gbs_block(h, chars, indices, key_nulls, values, value_nulls, n, 0, nullptr, stream);
}

int qk_gbs_process_block_where(qk_gbs h, const char* chars, const int64_t* indices, const uint64_t* key_nulls,
                               const void* values, const uint64_t* value_nulls, int32_t n, int32_t npreds,
                               const qk_predicate* preds, void* stream) {
    if (!h) return qk::fail(QK_ERROR, "qk_gbs_process_block_where: null table");
    return gbs_block(h, chars, indices, key_nulls, values, value_nulls, n, npreds, preds, stream);
}

int qk_gbs_get_results(qk_gbs h, char** out_chars, int64_t** out_indices, uint64_t** out_key_nulls,
                       void** out_values, uint64_t** out_value_nulls, int32_t* out_count, void* stream) {
    if (!h || !out_chars || !out_indices || !out_count) return qk::fail(QK_ERROR, "qk_gbs_get_results: null argument");
    *out_chars = nullptr;
    *out_indices = nullptr;
    if (out_key_nulls) *out_key_nulls = nullptr;
    void* keys[1] = {nullptr};
    uint64_t* kn[1] = {nullptr};
    int32_t g = 0;
    QK_TRY(qk_gb_get_results(h->inner, keys, kn, out_values, out_value_nulls, &g, stream));
    *out_count = g;
    if (g == 0) return QK_SUCCESS;
    hipStream_t s = qk::as_stream(stream);
    int64_t* lens = nullptr;
    int st = qk::pool_alloc((void**)&lens, (size_t)g * 8);
    const int64_t nb = ((int64_t)g + kScanRows - 1) / kScanRows;
    int64_t* sums = nullptr;
    if (st == QK_SUCCESS) st = qk::pool_alloc((void**)&sums, (size_t)nb * 8);
    char* chars = nullptr;
    if (st == QK_SUCCESS) {
        StrDict d = view(h);
        k_str_lengths<<<grid_for(g), 256, 0, s>>>(d, (const int64_t*)keys[0], kn[0], g, lens);
        k_scan64_sums<<<(unsigned)nb, kScanThreads, 0, s>>>(lens, g, sums);
        k_scan64_top<<<1, kScanThreads, 0, s>>>(sums, nb);
        k_scan64_apply<<<(unsigned)nb, kScanThreads, 0, s>>>(lens, g, sums);
        int64_t* hb = (int64_t*)h->host;
        if (hipMemcpyAsync(hb, lens + g - 1, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            st = qk::fail(QK_RUNTIME_ERROR, "string group by: reading the result size");
        if (st == QK_SUCCESS) st = qk::pool_alloc((void**)&chars, (size_t)std::max<int64_t>(hb[0], 1));
        if (st == QK_SUCCESS) {
            k_str_gather<<<grid_for(g), 256, 0, s>>>(d, (const int64_t*)keys[0], kn[0], g, lens, chars);
            if (hipGetLastError() != hipSuccess) st = qk::fail(QK_RUNTIME_ERROR, "string group by: gather launch");
        }
    }
    if (st == QK_SUCCESS) st = qk_stream_synchronize(stream);
    qk::pool_free(sums);
    qk::pool_free(keys[0]);
    if (st != QK_SUCCESS) {
        qk::pool_free(lens);
        qk::pool_free(chars);
        qk::pool_free(kn[0]);
        return st;
    }
    *out_chars = chars;
    *out_indices = lens;
    if (out_key_nulls) *out_key_nulls = kn[0];
    else qk::pool_free(kn[0]);
    return QK_SUCCESS;
}

int qk_gbs_group_count(qk_gbs h, int64_t* out, void* stream) {
    if (!h) return qk::fail(QK_ERROR, "qk_gbs_group_count: null table");
    return qk_gb_group_count(h->inner, out, stream);
}

}  // extern "C"
