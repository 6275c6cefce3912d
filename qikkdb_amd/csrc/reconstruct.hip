// reconstruct.hip - prefix sums and stream compaction.
//
// Counterparts of GPUReconstruct (QueryEngine/GPUCore/GPUReconstruct.cuh:29-71, :203-355,
// :486-624; GPUReconstruct.cu:8-43, :756-764).  The reference scans the mask with CUB and
// recomputes that scan for every compacted column; here the scan is our own single-pass
// decoupled look-back (lookback.hpp), and compaction computes in-tile offsets with wavefront
// ballots, so a column is compacted in one pass over mask + data without a materialised
// prefix array.  Null bits are compacted per wave with one OR-reduced word pair instead of
// one atomicOr per row.
#include <map>
#include <mutex>

#include "lookback.hpp"
#include "pred.hpp"

namespace qk {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;                        // int8 items per thread
constexpr int kScanTile = kScanBlock * kScanItems;    // 4096 rows per tile

constexpr int kCBlock = 256;
constexpr int kCSteps = 16;                           // rows per thread, one per 256-row step
constexpr int kCTile = kCBlock * kCSteps;             // 4096 rows per tile
constexpr int kMaxCols = 8;

// ---- inclusive / exclusive prefix sum of int8 values -> int32 ---------------------------
__global__ void __launch_bounds__(kScanBlock) k_scan_mask(const int8_t* __restrict__ mask,
                                                          int32_t* __restrict__ out, int64_t n,
                                                          int inclusive, uint64_t* status,
                                                          uint32_t* ticket, uint32_t epoch,
                                                          int64_t tiles) {
    __shared__ int64_t s_tile;
    __shared__ int32_t s_wave[kScanBlock / 64];
    __shared__ uint32_t s_excl;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (t == 0) s_tile = lb_ticket(ticket, tiles);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t r0 = tile * kScanTile + (int64_t)t * kScanItems;

    int32_t m[kScanItems];
    if (r0 + kScanItems <= n && (((uintptr_t)(mask + r0)) & 15u) == 0) {
        int4 v = *reinterpret_cast<const int4*>(mask + r0);
        const int8_t* b = reinterpret_cast<const int8_t*>(&v);
#pragma unroll
        for (int k = 0; k < kScanItems; k++) m[k] = b[k];
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) m[k] = (r0 + k < n) ? mask[r0 + k] : 0;
    }
    int32_t tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) tsum += m[k];

    const int32_t incl = wave_incl_scan(tsum, lane);
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    int32_t woff = 0, agg = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; w++) {
        if (w < wid) woff += s_wave[w];
        agg += s_wave[w];
    }
    if (wid == 0) {
        uint32_t e = lb_exclusive(status, epoch, tile, (uint32_t)agg);
        if (lane == 0) s_excl = e;
    }
    __syncthreads();
    int32_t run = (int32_t)s_excl + woff + (incl - tsum);
    int32_t o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (inclusive) {
            run += m[k];
            o[k] = run;
        } else {
            o[k] = run;
            run += m[k];
        }
    }
    if (r0 + kScanItems <= n && (((uintptr_t)(out + r0)) & 15u) == 0) {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; q++)
            reinterpret_cast<int4*>(out + r0)[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
            if (r0 + k < n) out[r0 + k] = o[k];
    }
}

// ---- reductions over a mask ------------------------------------------------------------
// SUM_VALUES: sum of the int8 values; otherwise count of non-zero bytes.
template <bool SUM_VALUES>
__global__ void __launch_bounds__(256) k_mask_reduce(const int8_t* __restrict__ mask, int64_t n,
                                                     unsigned long long* __restrict__ out) {
    __shared__ int64_t s_part[4];
    int64_t acc = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t vecs = n >> 4;
    const bool aligned = (((uintptr_t)mask) & 15u) == 0;
    if (aligned) {
        for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < vecs; v += stride) {
            int4 x = reinterpret_cast<const int4*>(mask)[v];
            const int8_t* b = reinterpret_cast<const int8_t*>(&x);
#pragma unroll
            for (int k = 0; k < 16; k++) acc += SUM_VALUES ? (int64_t)b[k] : (int64_t)(b[k] != 0);
        }
    }
    for (int64_t r = (aligned ? vecs * 16 : 0) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
         r < n; r += stride)
        acc += SUM_VALUES ? (int64_t)mask[r] : (int64_t)(mask[r] != 0);
    acc = wave_sum64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = s_part[0] + s_part[1] + s_part[2] + s_part[3];
        if (s) atomicAdd(out, (unsigned long long)s);
    }
}

// ---- stable compaction -------------------------------------------------------------------
struct CompactCols {
    int32_t ncols;
    int32_t esize[kMaxCols];
    const void* in[kMaxCols];
    const uint64_t* in_null[kMaxCols];
    void* out[kMaxCols];
    uint64_t* out_null[kMaxCols];
};

enum CompactMode { kByMask = 0, kByPreds = 1, kIndex32 = 2, kIndex64 = 3 };

__device__ __forceinline__ void copy_elem(void* out, const void* in, int64_t dst, int64_t src,
                                          int esize) {
    switch (esize) {
    case 1: ((int8_t*)out)[dst] = ((const int8_t*)in)[src]; break;
    case 2: ((int16_t*)out)[dst] = ((const int16_t*)in)[src]; break;
    case 4: ((int32_t*)out)[dst] = ((const int32_t*)in)[src]; break;
    default: ((int64_t*)out)[dst] = ((const int64_t*)in)[src]; break;
    }
}

// Row r of tile T sits at T*kCTile + k*kCBlock + threadIdx.x (k = step), so every load of a
// step is one coalesced 256-row span.  Rows are ordered by (step, thread); the in-tile offset
// of a kept row is (kept rows of earlier steps) + (kept rows of earlier waves in this step) +
// (kept lanes below it in its wave), all from ballots.
template <int MODE>
__global__ void __launch_bounds__(kCBlock) k_compact(const int8_t* __restrict__ mask, PredSet preds,
                                                     CompactCols cols, int64_t n,
                                                     uint64_t* status, uint32_t* ticket,
                                                     uint32_t epoch, int64_t tiles,
                                                     int32_t* __restrict__ total) {
    constexpr int W = kCBlock / 64;
    __shared__ int64_t s_tile;
    __shared__ int32_t s_cnt[kCSteps][W];
    __shared__ uint32_t s_excl;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (t == 0) s_tile = lb_ticket(ticket, tiles);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * kCTile;

    uint32_t keep = 0;  // bit k: row of step k is kept
    const uint64_t lt_mask = (1ull << lane) - 1;
    uint32_t pre[kCSteps];
#pragma unroll
    for (int k = 0; k < kCSteps; k++) {
        const int64_t r = base + (int64_t)k * kCBlock + t;
        bool sel = false;
        if (r < n) {
            if (MODE == kByPreds) {
                sel = true;
                for (int p = 0; p < preds.n && sel; p++) sel = eval_pred(preds.p[p], r);
            } else {
                sel = mask[r] != 0;
            }
        }
        const uint64_t b = __ballot(sel);
        pre[k] = (uint32_t)__popcll(b & lt_mask);
        if (lane == 0) s_cnt[k][wid] = (int32_t)__popcll(b);
        keep |= (sel ? 1u : 0u) << k;
    }
    __syncthreads();
    int32_t agg = 0;
    for (int k = 0; k < kCSteps; k++)
        for (int w = 0; w < W; w++) agg += s_cnt[k][w];
    if (wid == 0) {
        uint32_t e = lb_exclusive(status, epoch, tile, (uint32_t)agg);
        if (lane == 0) {
            s_excl = e;
            if (tile == tiles - 1) *total = (int32_t)(e + (uint32_t)agg);
        }
    }
    __syncthreads();
    int64_t step_base = (int64_t)s_excl;
#pragma unroll
    for (int k = 0; k < kCSteps; k++) {
        int64_t wave_base = step_base;
        for (int w = 0; w < wid; w++) wave_base += s_cnt[k][w];
        const bool sel = (keep >> k) & 1u;
        const int64_t r = base + (int64_t)k * kCBlock + t;
        const int64_t pos = wave_base + pre[k];
        if (MODE == kIndex32) {
            if (sel) ((int32_t*)cols.out[0])[pos] = (int32_t)r;
        } else if (MODE == kIndex64) {
            if (sel) ((int64_t*)cols.out[0])[pos] = r;
        } else {
            for (int c = 0; c < cols.ncols; c++) {
                if (sel) copy_elem(cols.out[c], cols.in[c], pos, r, cols.esize[c]);
                if (cols.out_null[c] && cols.in_null[c]) {
                    // null bits of this wave's kept rows land in [wave_base, wave_base + cnt)
                    const bool nb = sel && null_bit(cols.in_null[c], r);
                    const int off = (int)(pos - wave_base + (wave_base & 63));  // < 128
                    uint64_t lo = (nb && off < 64) ? (1ull << off) : 0ull;
                    uint64_t hi = (nb && off >= 64) ? (1ull << (off - 64)) : 0ull;
#pragma unroll
                    for (int d = 32; d >= 1; d >>= 1) {
                        lo |= __shfl_xor(lo, d, 64);
                        hi |= __shfl_xor(hi, d, 64);
                    }
                    if (lane == 0) {
                        uint64_t* dst = cols.out_null[c] + (wave_base >> 6);
                        if (lo) atomicOr((unsigned long long*)dst, (unsigned long long)lo);
                        if (hi) atomicOr((unsigned long long*)(dst + 1), (unsigned long long)hi);
                    }
                }
            }
        }
        int32_t step_total = 0;
        for (int w = 0; w < W; w++) step_total += s_cnt[k][w];
        step_base += step_total;
    }
}

// ---- one byte per row -> bit mask -------------------------------------------------------
__global__ void __launch_bounds__(256) k_compress_null_mask(const uint8_t* __restrict__ in,
                                                            uint64_t* __restrict__ out, int64_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t words = (n + 63) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < words; w += nwaves) {
        const int64_t r = (w << 6) + lane;
        const uint64_t b = __ballot(r < n && (in[r] & 1u));
        if (lane == 0) out[w] = b;
    }
}

int grid_stride_blocks(int64_t items, int per_block) {
    const DeviceInfo& d = device_info();
    int64_t g = (items + per_block - 1) / per_block;
    int64_t cap = (int64_t)(d.cus > 0 ? d.cus : 64) * 8;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// Launch k_compact over n rows; *count_host receives the kept count.
template <int MODE>
int launch_compact(const int8_t* mask, const PredSet& preds, const CompactCols& cols, int64_t n,
                   int32_t* count_host, hipStream_t s) {
    const int64_t tiles = (n + kCTile - 1) / kCTile;
    Scratch* sc = nullptr;
    QK_TRY(stream_scratch(s, lb_scratch_bytes(tiles) + 64, &sc));
    uint64_t* words = (uint64_t*)sc->ptr;
    uint32_t* ticket = (uint32_t*)words;
    uint64_t* status = words + kLbHeaderWords;
    int32_t* total = (int32_t*)(words + kLbHeaderWords + tiles);
    const uint32_t epoch = ++sc->epoch;
    k_compact<MODE><<<(unsigned)tiles, kCBlock, 0, s>>>(mask, preds, cols, n, status, ticket,
                                                        epoch, tiles, total);
    QK_HIP(hipGetLastError());
    if (count_host) {
        void* h = nullptr;
        QK_TRY(pinned_host(8, &h));
        QK_HIP(hipMemcpyAsync(h, total, 4, hipMemcpyDeviceToHost, s));
        QK_HIP(hipStreamSynchronize(s));
        *count_host = *(int32_t*)h;
    }
    return QK_SUCCESS;
}

// ---- prefix sum without allocation ----------------------------------------------------------
// qk_prefix_sum_noalloc keeps its look-back state (ticket word + one status word per
// 4096-row tile) in this fixed array instead of the per-stream scratch, so the call never
// allocates.  Each device holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
constexpr int64_t kNoAllocMaxRows = (int64_t)1 << 30;
constexpr int64_t kNoAllocTiles = kNoAllocMaxRows / kScanTile;
constexpr size_t kNoAllocWords = kLbHeaderWords + kNoAllocTiles + 8;
static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

struct NoAllocState {
    uint64_t* words = nullptr;  // device address of `scratch` on this device
    uint32_t epoch = 0;
};
std::mutex g_noalloc_mu;
std::map<int, NoAllocState> g_noalloc;

// Clears the ticket word before a launch (a kernel rather than a memset node, so the call can
// be captured into a graph).
__global__ void k_noalloc_arm(uint64_t* words) {
    if (threadIdx.x < kLbHeaderWords) words[threadIdx.x] = 0;
}

__global__ void k_noalloc_clear(uint64_t* words, size_t nwords) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) words[i] = 0;
}

// Resolves this device's copy of `scratch` and clears it, once per device (the first call on
// the device; it synchronises the device, so that call should not be made under capture).
int noalloc_state(NoAllocState** out) {
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    NoAllocState& st = g_noalloc[dev];
    if (!st.words) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QK_RUNTIME_ERROR, "prefix_sum_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        k_noalloc_clear<<<256, 256>>>((uint64_t*)p, kNoAllocWords);
        QK_HIP(hipGetLastError());
        QK_HIP(hipDeviceSynchronize());
        st.words = (uint64_t*)p;
    }
    *out = &st;
    return QK_SUCCESS;
}

}  // namespace

int prefix_noalloc_setup() {
    NoAllocState* st = nullptr;
    return noalloc_state(&st);
}

namespace {

// Count non-zero mask bytes (kept rows) on the device and read the count back.
int count_kept(const int8_t* mask, int64_t n, int64_t* out, hipStream_t s) {
    Scratch* sc = nullptr;
    QK_TRY(stream_scratch(s, 64, &sc));
    // the counter lives in its own scratch line (word 7) so look-back tickets are untouched
    unsigned long long* ctr = (unsigned long long*)sc->ptr + 7;
    QK_HIP(hipMemsetAsync(ctr, 0, 8, s));
    k_mask_reduce<false><<<grid_stride_blocks((n + 15) / 16, 256), 256, 0, s>>>(mask, n, ctr);
    QK_HIP(hipGetLastError());
    void* h = nullptr;
    QK_TRY(pinned_host(8, &h));
    QK_HIP(hipMemcpyAsync(h, ctr, 8, hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    *out = *(int64_t*)h;
    return QK_SUCCESS;
}

}  // namespace

}  // namespace qk

using namespace qk;

extern "C" {

int qk_prefix_sum(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive, void* stream) {
    QK_TRY(ensure_init());
    if (n <= 0) return QK_SUCCESS;
    if (!mask || !out) return fail(QK_ERROR, "prefix_sum: null buffer");
    hipStream_t s = as_stream(stream);
    const int64_t tiles = ((int64_t)n + kScanTile - 1) / kScanTile;
    Scratch* sc = nullptr;
    QK_TRY(stream_scratch(s, lb_scratch_bytes(tiles), &sc));
    uint64_t* words = (uint64_t*)sc->ptr;
    const uint32_t epoch = ++sc->epoch;
    k_scan_mask<<<(unsigned)tiles, kScanBlock, 0, s>>>(mask, out, n, inclusive ? 1 : 0,
                                                       words + kLbHeaderWords, (uint32_t*)words,
                                                       epoch, tiles);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

int qk_prefix_sum_noalloc(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive, void* stream) {
    QK_TRY(ensure_init());
    if (n <= 0) return QK_SUCCESS;
    if (!mask || !out) return fail(QK_ERROR, "prefix_sum_noalloc: null buffer");
    if ((int64_t)n > kNoAllocMaxRows)
        return fail(QK_ERROR, "prefix_sum_noalloc: %d rows exceed the fixed scratch's maximum of %lld rows",
                    n, (long long)kNoAllocMaxRows);
    NoAllocState* st = nullptr;
    QK_TRY(noalloc_state(&st));
    uint32_t epoch;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        epoch = ++st->epoch;
    }
    hipStream_t s = as_stream(stream);
    const int64_t tiles = ((int64_t)n + kScanTile - 1) / kScanTile;
    k_noalloc_arm<<<1, 64, 0, s>>>(st->words);
    QK_HIP(hipGetLastError());
    k_scan_mask<<<(unsigned)tiles, kScanBlock, 0, s>>>(mask, out, n, inclusive ? 1 : 0,
                                                       st->words + kLbHeaderWords, (uint32_t*)st->words,
                                                       epoch, tiles);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

int qk_mask_sum(const int8_t* mask, int32_t n, int64_t* out_host, void* stream) {
    QK_TRY(ensure_init());
    if (!out_host) return fail(QK_ERROR, "mask_sum: null output");
    *out_host = 0;
    if (n <= 0) return QK_SUCCESS;
    if (!mask) return fail(QK_ERROR, "mask_sum: null mask");
    hipStream_t s = as_stream(stream);
    Scratch* sc = nullptr;
    QK_TRY(stream_scratch(s, 64, &sc));
    unsigned long long* ctr = (unsigned long long*)sc->ptr + 7;
    QK_HIP(hipMemsetAsync(ctr, 0, 8, s));
    k_mask_reduce<true><<<grid_stride_blocks(((int64_t)n + 15) / 16, 256), 256, 0, s>>>(mask, n, ctr);
    QK_HIP(hipGetLastError());
    void* h = nullptr;
    QK_TRY(pinned_host(8, &h));
    QK_HIP(hipMemcpyAsync(h, ctr, 8, hipMemcpyDeviceToHost, s));
    QK_HIP(hipStreamSynchronize(s));
    *out_host = *(int64_t*)h;
    return QK_SUCCESS;
}

int qk_reconstruct_col_keep(int32_t elem_size, const void* in, const int8_t* mask, int32_t n,
                            const uint64_t* in_null, void** out, int32_t* out_count,
                            uint64_t** out_null, void* stream) {
    QK_TRY(ensure_init());
    if (!out || !out_count) return fail(QK_ERROR, "reconstruct_col_keep: null output");
    if (elem_size != 1 && elem_size != 2 && elem_size != 4 && elem_size != 8)
        return fail(QK_ERROR, "reconstruct_col_keep: element size %d unsupported", elem_size);
    if (n <= 0) {  // GPUReconstruct.cuh:330-338
        *out = nullptr;
        *out_count = 0;
        if (out_null) *out_null = nullptr;
        return QK_SUCCESS;
    }
    if (!mask) {  // aliasing, no ownership transfer (GPUReconstruct.cuh:320-328)
        *out = const_cast<void*>(in);
        *out_count = n;
        if (out_null) *out_null = const_cast<uint64_t*>(in_null);
        return QK_SUCCESS;
    }
    if (!in) return fail(QK_ERROR, "reconstruct_col_keep: null input");
    hipStream_t s = as_stream(stream);
    int64_t kept = 0;
    QK_TRY(count_kept(mask, n, &kept, s));
    *out_count = (int32_t)kept;
    if (kept == 0) {  // GPUReconstruct.cuh:311-318
        *out = nullptr;
        if (out_null) *out_null = nullptr;
        return QK_SUCCESS;
    }
    void* o = nullptr;
    uint64_t* on = nullptr;
    QK_TRY(pool_alloc(&o, (size_t)kept * elem_size));
    const bool want_null = out_null && in_null;
    if (want_null) {
        int st = pool_alloc((void**)&on, (size_t)((kept + 63) / 64) * 8);
        if (st != QK_SUCCESS) {
            pool_free(o);
            return st;
        }
        QK_HIP(hipMemsetAsync(on, 0, (size_t)((kept + 63) / 64) * 8, s));
    }
    CompactCols cols{};
    cols.ncols = 1;
    cols.esize[0] = elem_size;
    cols.in[0] = in;
    cols.in_null[0] = want_null ? in_null : nullptr;
    cols.out[0] = o;
    cols.out_null[0] = on;
    PredSet none{};
    int st = launch_compact<kByMask>(mask, none, cols, n, nullptr, s);
    if (st != QK_SUCCESS) {
        pool_free(o);
        pool_free(on);
        return st;
    }
    *out = o;
    if (out_null) *out_null = on;
    return QK_SUCCESS;
}

int qk_generate_indexes_keep(const int8_t* mask, int32_t n, int32_t index_size, void** out,
                             int32_t* out_count, void* stream) {
    QK_TRY(ensure_init());
    if (!out || !out_count) return fail(QK_ERROR, "generate_indexes: null output");
    if (!mask) return fail(QK_ERROR, "inMask cannot be nullptr in GenerateIndexes");  // :558-562
    if (index_size != 4 && index_size != 8)
        return fail(QK_ERROR, "generate_indexes: index size must be 4 or 8");
    *out = nullptr;
    *out_count = 0;
    if (n <= 0) return QK_SUCCESS;
    hipStream_t s = as_stream(stream);
    int64_t kept = 0;
    QK_TRY(count_kept(mask, n, &kept, s));
    *out_count = (int32_t)kept;
    if (kept == 0) return QK_SUCCESS;
    void* o = nullptr;
    QK_TRY(pool_alloc(&o, (size_t)kept * index_size));
    CompactCols cols{};
    cols.ncols = 1;
    cols.out[0] = o;
    PredSet none{};
    int st = index_size == 4 ? launch_compact<kIndex32>(mask, none, cols, n, nullptr, s)
                             : launch_compact<kIndex64>(mask, none, cols, n, nullptr, s);
    if (st != QK_SUCCESS) {
        pool_free(o);
        return st;
    }
    *out = o;
    return QK_SUCCESS;
}

int qk_compress_null_mask(const uint8_t* in, int32_t n, uint64_t** out, void* stream) {
    QK_TRY(ensure_init());
    if (!out) return fail(QK_ERROR, "compress_null_mask: null output");
    *out = nullptr;
    if (n <= 0) return QK_SUCCESS;
    if (!in) return fail(QK_ERROR, "compress_null_mask: null input");
    const int64_t words = ((int64_t)n + 63) / 64;
    uint64_t* o = nullptr;
    QK_TRY(pool_alloc((void**)&o, (size_t)words * 8));
    hipStream_t s = as_stream(stream);
    k_compress_null_mask<<<grid_stride_blocks(words, 4), 256, 0, s>>>(in, o, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        pool_free(o);
        return fail(QK_RUNTIME_ERROR, "compress_null_mask launch: %s", hipGetErrorString(e));
    }
    *out = o;
    return QK_SUCCESS;
}

int qk_where_compact(int32_t npreds, const qk_predicate* preds, int32_t ncols,
                     const int32_t* elem_sizes, const void* const* cols, const uint64_t* const* col_nulls,
                     int32_t n, void* const* outs, uint64_t* const* out_nulls, int32_t* out_count,
                     void* stream) {
    QK_TRY(ensure_init());
    if (!out_count) return fail(QK_ERROR, "where_compact: null count output");
    *out_count = 0;
    if (ncols < 0 || ncols > kMaxCols) return fail(QK_ERROR, "where_compact: at most %d columns", kMaxCols);
    PredSet ps{};
    QK_TRY(check_preds(npreds, preds, &ps));
    if (n <= 0) return QK_SUCCESS;
    CompactCols cc{};
    cc.ncols = ncols;
    for (int c = 0; c < ncols; c++) {
        const int es = elem_sizes[c];
        if (es != 1 && es != 2 && es != 4 && es != 8)
            return fail(QK_ERROR, "where_compact: element size %d unsupported", es);
        if (!cols[c] || !outs[c]) return fail(QK_ERROR, "where_compact: null column %d", c);
        cc.esize[c] = es;
        cc.in[c] = cols[c];
        cc.in_null[c] = col_nulls ? col_nulls[c] : nullptr;
        cc.out[c] = outs[c];
        cc.out_null[c] = (out_nulls && cc.in_null[c]) ? out_nulls[c] : nullptr;
    }
    hipStream_t s = as_stream(stream);
    for (int c = 0; c < ncols; c++)
        if (cc.out_null[c]) QK_HIP(hipMemsetAsync(cc.out_null[c], 0, (size_t)(((int64_t)n + 63) / 64) * 8, s));
    return launch_compact<kByPreds>(nullptr, ps, cc, n, out_count, s);
}

}  // extern "C"
