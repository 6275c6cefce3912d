// twophase.hip - two-phase partitioned GROUP BY for single-word keys (see twophase.hpp).
//
// Heavy hitters: a strided sample of the first block is counted (per workgroup in LDS, then
// merged into a device hash table); the keys seen at least `t` times (t chosen from the count
// histogram so that at most kmax keys qualify) become the heavy set, sorted by partition.
// All of it runs on the device.
#include <algorithm>
#include <cstring>

#include "twophase.hpp"

namespace qk {

namespace {

constexpr int kP1Steps = 4;   // rows per thread per tile in phase 1
constexpr int kP2Rows = 6;    // cold rows per thread per pass in phase 2
constexpr int kBlock = 1024;
constexpr int kBucket = 4;    // slots per bucket of the phase-1 heavy table
constexpr int kSampleLds = 2048;
// threads per sampling workgroup: each flushes kSampleLds / kSampleBlock slots of its LDS
// table into the device table, one probe at a time
constexpr int kSampleBlock = 1024;

// A cold row in its region: the key and the value packed into one record of 32-bit words
// (8, 12 or 16 bytes), so a workgroup fills one line per partition at a time instead of one
// line per column; fewer lines are half-written when the L2 evicts them.
template <typename VT>
struct RecOf {
    static constexpr int words = 2 + (VTraits<VT>::has ? (int)(sizeof(VT) / 4) : 0);
    struct alignas(4) Rec { uint32_t w[words]; };
};

template <typename VT, typename VV>
__device__ __forceinline__ void put_rec(uint32_t* base, int64_t at, uint64_t key, const VV& v) {
    using R = RecOf<VT>;
    typename R::Rec r;
    r.w[0] = (uint32_t)key;
    r.w[1] = (uint32_t)(key >> 32);
    if constexpr (VTraits<VT>::has) __builtin_memcpy(&r.w[2], &v, sizeof(VT));
    *reinterpret_cast<typename R::Rec*>(base + at * R::words) = r;
}

template <typename VT, typename VV>
__device__ __forceinline__ void get_rec(const uint32_t* base, int64_t at, uint64_t& key, VV& v) {
    using R = RecOf<VT>;
    const typename R::Rec r = *reinterpret_cast<const typename R::Rec*>(base + at * R::words);
    key = (uint64_t)r.w[0] | ((uint64_t)r.w[1] << 32);
    if constexpr (VTraits<VT>::has) __builtin_memcpy(&v, &r.w[2], sizeof(VT)); else v = 0;
}

// VV is VT, or a placeholder type when the table reads no value column.  MODE: 0 no WHERE,
// 1 row mask, 2 predicate list, 3 folded interval on the value column, 4 folded interval on
// another integer column, whose value (cv) and NULL bit (cn) the caller loaded.
template <typename VT, int MODE, typename VV>
__device__ __forceinline__ bool row_passes(const PredSet& preds, uint32_t vpred, const int8_t* mask,
                                           int64_t r, const VV& v, bool vn, int64_t cv = 0, bool cn = false) {
    if (MODE == 1) return mask[r] != 0;
    // & rather than &&: both sides are cheap and side-effect free, and a short circuit would
    // become a divergent branch per row
    if constexpr (MODE == 3 && VTraits<VT>::has) return !vn & in_range<VT>(preds.vr, v);
    if (MODE == 4) return !cn & col_in_range(preds, cv);
    if (MODE == 2) {
        for (int p = 0; p < preds.n; p++) {
            const qk_predicate& q = preds.p[p];
            bool pass;
            if constexpr (VTraits<VT>::has) {
                if ((vpred >> p) & 1u)
                    pass = !vn && pred_on_value(q, v, r);
                else
                    pass = eval_pred(q, r);
            } else {
                pass = eval_pred(q, r);
            }
            if (!pass) return false;
        }
    }
    return true;
}

// Exclusive prefix sum over the threads of a workgroup (blockDim a multiple of 64, <= 1024);
// *total receives the sum.  `wsum` is LDS for 16 values.  Every thread must call it.
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* wsum, int32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int32_t incl = wave_incl_scan(v, lane);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int32_t before = 0, all = 0;
    for (int w = 0; w < nw; w++) {
        if (w < wid) before += wsum[w];
        all += wsum[w];
    }
    *total = all;
    __syncthreads();
    return before + incl - v;
}

// ---- sampling --------------------------------------------------------------------------

__global__ void k_tp_clear_sample(uint64_t* keys, uint32_t* cnt, uint64_t slots, uint32_t* hist,
                                  int32_t* meta) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += stride) {
        keys[s] = kEmpty;
        cnt[s] = 0;
    }
    if (blockIdx.x == 0) {
        for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) hist[b] = 0;
        if (threadIdx.x == 0) meta[0] = 0;
    }
}

// Add `c` occurrences of key k to the device sample table.
__device__ __forceinline__ void sample_add(uint64_t* skeys, uint32_t* scnt, uint64_t smask,
                                           uint64_t k, uint32_t c) {
    uint64_t s = hash64(k) & smask;
    for (uint64_t p = 0; p <= smask; p++) {
        uint64_t cur = *(volatile uint64_t*)(skeys + s);
        if (cur == kEmpty) {
            cur = atomicCAS((unsigned long long*)(skeys + s), (unsigned long long)kEmpty,
                            (unsigned long long)k);
            if (cur == kEmpty) cur = k;
        }
        if (cur == k) {
            atomicAdd(scnt + s, c);
            return;
        }
        s = (s + 1) & smask;
    }
}

// Each workgroup counts a contiguous range of samples in LDS first, so that a hot key costs
// one device atomic per workgroup instead of one per sample.
template <typename VT, int MODE>
__global__ void __launch_bounds__(kSampleBlock) k_tp_sample(KeySpec ks, KeyCols kc, const VT* __restrict__ vals,
                                                   const uint64_t* __restrict__ vnull,
                                                   const int8_t* __restrict__ mask, PredSet preds,
                                                   uint32_t vpred, int64_t n, int64_t stride,
                                                   int64_t nsamples, uint64_t* skeys,
                                                   uint32_t* scnt, uint64_t smask) {
    __shared__ uint64_t lk[kSampleLds];
    __shared__ uint32_t lc[kSampleLds];
    for (int s = threadIdx.x; s < kSampleLds; s += blockDim.x) {
        lk[s] = kEmpty;
        lc[s] = 0;
    }
    __syncthreads();
    const int64_t per = (nsamples + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * per, i1 = min(nsamples, i0 + per);
    // kBatch samples per thread at a time, their loads issued before any is used (the samples
    // are rows far apart, so each is its own memory round trip)
    constexpr int kBatch = 4;
    using VV = typename std::conditional<VTraits<VT>::has, VT, int>::type;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kBatch * (int64_t)blockDim.x) {
        VV v[kBatch];
        bool vn[kBatch], cn[kBatch], kn[kBatch], ok[kBatch];
        int64_t cv[kBatch], rr[kBatch];
        uint64_t k[kBatch][1];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            // runs of 16 consecutive rows, `stride` sample rows apart, so that the loads of a
            // wavefront touch few lines
            const int64_t iu = i + (int64_t)u * blockDim.x;
            const int64_t r = (iu >> 4) * (stride << 4) + (iu & 15);
            ok[u] = iu < i1 && r < n;
            rr[u] = ok[u] ? r : 0;
            v[u] = VV{};
            if constexpr (VTraits<VT>::has) v[u] = vals[rr[u]];
            vn[u] = vnull && null_bit(vnull, rr[u]);
            cv[u] = 0;
            cn[u] = false;
            if (MODE == 4) {
                cv[u] = range_col(preds, rr[u]);
                cn[u] = preds.rnull && null_bit(preds.rnull, rr[u]);
            }
            load_key<1>(ks, kc, rr[u], k[u], kn[u]);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
        if (!ok[u] || !row_passes<VT, MODE>(preds, vpred, mask, rr[u], v[u], vn[u], cv[u], cn[u])) continue;
        if (kn[u] || k[u][0] == kEmpty) continue;  // those keys have side slots anyway
        const uint64_t key = k[u][0];
        uint32_t s = (uint32_t)(hash64(key) >> 53) & (kSampleLds - 1);
        bool done = false;
        for (int p = 0; p < 16 && !done; p++) {
            unsigned long long old = atomicCAS((unsigned long long*)&lk[s], (unsigned long long)kEmpty,
                                               (unsigned long long)key);
            if (old == kEmpty || old == key) {
                atomicAdd(&lc[s], 1u);
                done = true;
            }
            s = (s + 1) & (kSampleLds - 1);
        }
        if (!done) sample_add(skeys, scnt, smask, key, 1u);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kSampleLds; s += blockDim.x)
        if (lc[s]) sample_add(skeys, scnt, smask, lk[s], lc[s]);
}

// Histogram of the sample counts (clamped to kHistBins - 1): per workgroup in LDS, then one
// device atomic per non-empty bin and workgroup, so few large workgroups keep the device
// atomics on the hot low bins few.
constexpr int kHistBlock = 1024;
constexpr int kHistGrid = 64;
__global__ void __launch_bounds__(kHistBlock) k_tp_hist(const uint64_t* __restrict__ skeys,
                                                 const uint32_t* __restrict__ scnt, uint64_t slots,
                                                 uint32_t* hist) {
    __shared__ uint32_t lh[kHistBins];
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += stride)
        if (skeys[s] != kEmpty) atomicAdd(&lh[min(scnt[s], (uint32_t)(kHistBins - 1))], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x)
        if (lh[b]) atomicAdd(hist + b, lh[b]);
}

// Smallest count t >= 2 such that at most kmax sampled keys were seen t times or more
// (kHistBins threads: thread i holds bin kHistBins-1-i, so a prefix sum is a suffix sum).
// When the keys seen t times or more hold less than 1/16 of the sampled rows, there is no
// heavy set (t = INT32_MAX): a heavy lookup costs every row a little, and so few heavy rows do
// not repay it.
__global__ void __launch_bounds__(kHistBins) k_tp_threshold(const uint32_t* hist, int32_t kmax,
                                                            int32_t* meta) {
    __shared__ int32_t wsum[16];
    __shared__ int32_t best;
    const int i = threadIdx.x, b = kHistBins - 1 - i;
    if (i == 0) best = kHistBins;
    const int32_t v = (int32_t)min(hist[b], (uint32_t)0x3fffffff);
    int32_t total;
    const int32_t suffix = block_excl_scan(v, wsum, &total) + v;  // keys seen >= b times
    if (b >= 2 && suffix <= kmax) atomicMin(&best, b);
    __syncthreads();
    const int32_t t = best;
    // sampled rows in all and of the keys seen >= t times (the last bin counts its keys
    // kHistBins - 1 times each, at most 2^18 rows in all)
    const int64_t rows_b = (int64_t)b * v;
    const int32_t rb = (int32_t)min<int64_t>(rows_b, 0x3fffffff);
    int32_t all_rows, heavy_rows;
    (void)block_excl_scan(rb, wsum, &all_rows);
    (void)block_excl_scan(b >= t ? rb : 0, wsum, &heavy_rows);
    if (i == 0) {
        meta[1] = (int64_t)heavy_rows * 16 < (int64_t)all_rows ? 0x7fffffff : t;
        meta[0] = 0;
    }
}

// Keys seen at least meta[1] times, appended to cand / cand_cnt (meta[0] counts them).  Each
// workgroup collects the candidates of its range of slots in LDS and reserves their place with
// one atomic, so the counter sees one atomic per workgroup (at most kmax candidates in all).
constexpr int kCollectBlock = 1024;
constexpr int kCollectGrid = 64;
__global__ void __launch_bounds__(kCollectBlock) k_tp_collect(const uint64_t* __restrict__ skeys,
                                                              const uint32_t* __restrict__ scnt, uint64_t slots,
                                                              int32_t kmax, int32_t* meta, uint64_t* cand,
                                                              uint32_t* cand_cnt) {
    __shared__ uint64_t lkey[4096];
    __shared__ uint32_t lcnt[4096];
    __shared__ int32_t ln, lbase;
    const uint32_t t = (uint32_t)meta[1];
    if (threadIdx.x == 0) ln = 0;
    __syncthreads();
    const uint64_t per = (slots + gridDim.x - 1) / gridDim.x;
    const uint64_t s0 = (uint64_t)blockIdx.x * per, s1 = min(slots, s0 + per);
    for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
        const uint64_t key = skeys[s];
        const uint32_t c = scnt[s];
        if (key != kEmpty && c >= t) {
            const int32_t i = atomicAdd(&ln, 1);
            if (i < 4096) {
                lkey[i] = key;
                lcnt[i] = c;
            }
        }
    }
    __syncthreads();
    const int32_t n = min(ln, min(kmax, 4096));
    if (threadIdx.x == 0) lbase = n ? atomicAdd(meta, n) : 0;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t at = lbase + i;
        if (at < kmax) {
            cand[at] = lkey[i];
            cand_cnt[at] = lcnt[i];
        }
    }
}

// ---- heavy set order and heavy table image ---------------------------------------------------

// The bucket of a key in the heavy table: a multiplicative hash of the key's two halves folded
// together (Knuth's golden-ratio multiplier; the high bits of the product pick the bucket).
// The table only needs keys spread over the buckets, so this stands in for the full 64-bit
// hash on the path every row takes.
__device__ __forceinline__ uint32_t bucket_of(uint64_t k, uint32_t nb) {
    const uint32_t f = ((uint32_t)k ^ (uint32_t)(k >> 32)) * 0x9e3779b1u;
    return __umulhi(f, nb);
}

// One workgroup, once per heavy set:
// 1. orders the heavy candidates by partition (heavy[], heavy_cnt[]; meta[2 + p] = first
//    index of partition p, meta[2 + P] = K, meta[0] = K);
// 2. places each heavy key in a free slot of its bucket of an H-slot table, the most frequent
//    keys first (one round per power of two of the sample count), so that the keys carrying
//    most rows always find a slot.  A key that finds its bucket full stays cold (its rows then
//    go through its partition like any other key's).  The table's keys go to image[H] (kEmpty
//    in free slots) and each heavy key's slot to slot[k] (-1 when not placed); every phase-1
//    workgroup starts from this image.
// Dynamic LDS: H keys of the table + kmax keys and counts of the ordered set.
__global__ void __launch_bounds__(1024) k_tp_order(const uint64_t* __restrict__ cand,
                                                   const uint32_t* __restrict__ cand_cnt, int32_t kmax,
                                                   int32_t P, int32_t part_shift, int32_t* meta,
                                                   uint64_t* __restrict__ heavy, uint32_t* __restrict__ heavy_cnt,
                                                   int32_t H, uint64_t* __restrict__ image,
                                                   int32_t* __restrict__ slot) {
    extern __shared__ __align__(16) uint64_t lds[];
    __shared__ int32_t cnt[kMaxRecParts];
    __shared__ int32_t cur[kMaxRecParts];
    __shared__ int32_t wsum[16];
    uint64_t* lkey = lds;                                       // [H]
    uint64_t* okey = lds + H;                                   // [kmax] ordered heavy keys
    uint32_t* ocnt = reinterpret_cast<uint32_t*>(okey + kmax);  // [kmax] their sample counts
    const int t = threadIdx.x, bd = blockDim.x;
    const int32_t K = min(meta[0], kmax);
    for (int p = t; p < P; p += bd) cnt[p] = 0;
    for (int i = t; i < H; i += bd) lkey[i] = kEmpty;
    __syncthreads();
    for (int i = t; i < K; i += bd) {
        const uint64_t h = hash64(cand[i]);
        const int p = part_shift >= 64 ? 0 : (int)(h >> part_shift);
        atomicAdd(&cnt[p], 1);
    }
    __syncthreads();
    int32_t total;
    const int32_t start = block_excl_scan(t < P ? cnt[t] : 0, wsum, &total);
    if (t < P) {
        cur[t] = start;
        meta[2 + t] = start;
    }
    if (t == 0) {
        meta[2 + P] = total;
        meta[0] = K;
    }
    __syncthreads();
    for (int i = t; i < K; i += bd) {
        const uint64_t key = cand[i];
        const uint64_t h = hash64(key);
        const int p = part_shift >= 64 ? 0 : (int)(h >> part_shift);
        const int32_t at = atomicAdd(&cur[p], 1);
        heavy[at] = key;
        heavy_cnt[at] = cand_cnt[i];
        okey[at] = key;
        ocnt[at] = cand_cnt[i];
    }
    __syncthreads();

    const uint32_t nb = (uint32_t)(H / kBucket);
    // kmax <= 4 * blockDim, so each thread holds at most kPlace heavy keys
    constexpr int kPlace = 4;
    static_assert(kPlace * 1024 >= 4096, "heavy keys per thread");
    uint64_t pkey[kPlace];
    int pcls[kPlace];
#pragma unroll
    for (int i = 0; i < kPlace; i++) {
        const int k = t + i * bd;
        pcls[i] = -1;
        pkey[i] = 0;
        if (k < K) {
            pkey[i] = okey[k];
            pcls[i] = 31 - __builtin_clz(max(ocnt[k], 1u));
        }
    }
    // the classes present, so that only they take a barrier each
    __shared__ uint32_t s_classes;
    if (t == 0) s_classes = 0;
    __syncthreads();
    uint32_t my_classes = 0;
#pragma unroll
    for (int i = 0; i < kPlace; i++) my_classes |= pcls[i] >= 0 ? 1u << pcls[i] : 0u;
    if (my_classes) atomicOr(&s_classes, my_classes);
    __syncthreads();
    const uint32_t classes = s_classes;
    for (int cls = 31; cls >= 0; cls--) {
        if (!((classes >> cls) & 1u)) continue;
        __syncthreads();  // the keys of the classes above are placed
#pragma unroll
        for (int i = 0; i < kPlace; i++) {
            if (pcls[i] != cls) continue;
            const uint32_t bb = bucket_of(pkey[i], nb);
            int placed = -1;
            for (int j = 0; j < kBucket && placed < 0; j++) {
                const unsigned long long old = atomicCAS((unsigned long long*)&lkey[bb * kBucket + j],
                                                         (unsigned long long)kEmpty, (unsigned long long)pkey[i]);
                if (old == kEmpty) placed = (int)(bb * kBucket + j);
            }
            slot[t + i * bd] = placed;
        }
    }
    __syncthreads();
    for (int i = t; i < H; i += bd) image[i] = lkey[i];
}

// ---- phase 1 -------------------------------------------------------------------------------

struct P1Args {
    const uint64_t* image;      // [H] keys of the heavy table (k_tp_order)
    const int32_t* slot;        // [kmax] slot of each heavy key in it, -1 if not placed
    const int32_t* meta;
    int32_t kmax, H, P, cap;
    uint32_t* p_cnt;
    uint32_t* p_rows;
    uint64_t* p_sum;
    uint64_t* p_min;
    uint64_t* p_max;
    uint32_t* r_rec;       // cold-row records [W][P][cap] (RecOf<VT>)
    uint32_t* r_count;
    const uint64_t* zero;  // a zero word (stands in for an absent null bitmap)
    int32_t part_shift;    // partition of a key hash: h >> part_shift (0 when >= 64)
    int32_t accumulate;    // regions and partials hold rows of earlier blocks
};

// Cold rows wait in a ring of kQueue entries per wavefront until 64 of them are there, then
// the whole wavefront partitions and writes those 64 at once.
constexpr int kQueue = 128;

// LDS of phase 1: H hashed slots in H/kBucket buckets + 2 side slots for the partial arrays,
// the cold-row rings of the workgroup's wavefronts (key + value of vbytes), the partition
// cursors.
__host__ __device__ inline size_t p1_lds_bytes(int32_t H, int32_t P, int32_t vbytes, uint32_t fields) {
    const size_t T = (size_t)H + 2;
    size_t b = (size_t)H * 8;                                  // keys
    b += T * 8 * (((fields & F_SUM) ? 1 : 0) + ((fields & F_MIN) ? 1 : 0) + ((fields & F_MAX) ? 1 : 0));
    b += T * 4 * 2;                                            // cnt, rows
    b += (size_t)(kBlock / 64) * kQueue * (8 + (size_t)vbytes); // cold-row rings
    b += (size_t)P * 4 + 16;                                   // cursors, counters
    return b;
}

// Slot of key k in the heavy table, or -1 (one bucket of kBucket keys, read 128 bits at a time).
__device__ __forceinline__ int heavy_find(const uint64_t* lkey, uint32_t nb, uint64_t k) {
    static_assert(kBucket == 2 || kBucket == 4, "bucket width");
    const uint32_t b = bucket_of(k, nb);
    const ulonglong2* p = reinterpret_cast<const ulonglong2*>(lkey + (size_t)b * kBucket);
    const int base = (int)(b * kBucket);
    const ulonglong2 a0 = p[0];
    int s = -1;
    s = a0.x == k ? base : s;
    s = a0.y == k ? base + 1 : s;
    if constexpr (kBucket == 4) {
        const ulonglong2 a1 = p[1];
        s = a1.x == k ? base + 2 : s;
        s = a1.y == k ? base + 3 : s;
    }
    return s;
}

// KT is the element type of a single integer key column, KeyPairI32 for two non-NULL int32
// key columns packed into one word (the first in the low half), or void for any key layout
// (read through the key spec's runtime switch).  With a typed key, the loads of a whole tile
// are issued before any of them is used.
struct KeyPairI32 {};
template <typename KT>
constexpr bool kPairKey = std::is_same<KT, KeyPairI32>::value;
template <typename KT>
constexpr bool kTypedKey = !std::is_same<KT, void>::value && !kPairKey<KT>;

// The global table is read through a pointer: only the rare atomic path needs it, and kernel
// arguments kept in scalar registers for the whole loop would spill.
//
// FP selects the aggregate states the LDS table keeps: kFieldsSum fixes them at sum + count
// (SUM, AVG, COUNT; no MIN/MAX) so that the loop carries no per-field tests; kFieldsAny reads
// them from `fields`.
constexpr int kFieldsAny = 0, kFieldsSum = 1;

// The phase-1 LDS of one workgroup, carved as p1_lds_bytes lays it out.
template <typename VV>
struct P1Lds {
    uint64_t* lkey;
    uint64_t* lsum;
    uint64_t* lmin;
    uint64_t* lmax;
    uint64_t* qkey;   // this wavefront's cold-row ring: keys
    VV* qval;         // and values
    uint32_t* lcnt;   // rows with a non-NULL value, per slot
    uint32_t* lnul;   // rows with a NULL value, per slot
    uint32_t* cursor; // per partition: records in this workgroup's region
    uint32_t* nnew;   // keys the atomic path added to the global table
};

// Lays out the LDS, copies the heavy table's image in, clears the states and loads the
// region cursors (continuing earlier blocks' regions when accumulating).  Ends with a barrier.
template <typename VT, typename VV>
__device__ __forceinline__ P1Lds<VV> p1_begin(uint64_t* lds, const P1Args& a, bool f_sum, bool f_min,
                                              bool f_max) {
    constexpr bool HAS_V = VTraits<VT>::has;
    const int H = a.H, T = H + 2, bd = blockDim.x, tid = threadIdx.x, wave = tid >> 6;
    P1Lds<VV> L;
    L.lkey = lds;
    uint64_t* p = L.lkey + H;
    L.lsum = f_sum ? p : nullptr;
    p += f_sum ? T : 0;
    L.lmin = f_min ? p : nullptr;
    p += f_min ? T : 0;
    L.lmax = f_max ? p : nullptr;
    p += f_max ? T : 0;
    // the cold-row rings of the wavefronts: keys [waves][kQueue], then values
    L.qkey = p + (size_t)wave * kQueue;
    p += (size_t)(kBlock / 64) * kQueue;
    L.qval = reinterpret_cast<VV*>(p) + (size_t)wave * kQueue;
    p = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(p) + (size_t)(kBlock / 64) * kQueue * (HAS_V ? sizeof(VT) : 0));
    L.lcnt = (uint32_t*)p;
    L.lnul = L.lcnt + T;
    L.cursor = L.lnul + T;
    L.nnew = L.cursor + a.P;
    // the heavy table's keys from the image (two keys per 16-byte load), states cleared
    {
        const ulonglong2* img = reinterpret_cast<const ulonglong2*>(a.image);
        ulonglong2* dst = reinterpret_cast<ulonglong2*>(L.lkey);
        for (int i = tid; i < H / 2; i += bd) dst[i] = img[i];
    }
    for (int s = tid; s < T; s += bd) {
        L.lcnt[s] = 0;
        L.lnul[s] = 0;
        if (L.lsum) L.lsum[s] = 0;
        if (L.lmin) L.lmin[s] = ~0ull;
        if (L.lmax) L.lmax[s] = 0;
    }
    const int64_t region0 = (int64_t)blockIdx.x * a.P;
    // regions continue where the previous block left them until phase 2 consumes them
    for (int q = tid; q < a.P; q += bd) L.cursor[q] = a.accumulate ? a.r_count[region0 + q] : 0u;
    if (tid == 0) L.nnew[0] = 0;
    __syncthreads();
    return L;
}

// Writes the workgroup's heavy partials (indexed by heavy key; zeros for a key the image could
// not place), added to what earlier blocks left when phase 2 has not consumed it yet, the
// region fill counts and the count of keys the atomic path added.  Starts with a barrier.
template <bool REAL, typename VV>
__device__ __forceinline__ void p1_end(const P1Lds<VV>& L, const P1Args& a, const GView* gp) {
    __syncthreads();
    const int H = a.H, bd = blockDim.x, tid = threadIdx.x;
    const int32_t K = min(a.meta[0], a.kmax);
    const int64_t pbase = (int64_t)blockIdx.x * (a.kmax + 2);
    for (int k = tid; k < K + 2; k += bd) {
        const int s = k < K ? a.slot[k] : H + (k - K);
        const int64_t at = pbase + (k < K ? k : a.kmax + (k - K));
        uint32_t cnt = s >= 0 ? L.lcnt[s] : 0u, rows = s >= 0 ? L.lcnt[s] + L.lnul[s] : 0u;
        uint64_t sum = (L.lsum && s >= 0) ? L.lsum[s] : 0ull;
        uint64_t mn = (L.lmin && s >= 0) ? L.lmin[s] : ~0ull, mx = (L.lmax && s >= 0) ? L.lmax[s] : 0ull;
        if (a.accumulate) {
            rows += a.p_rows[at];
            cnt += a.p_cnt[at];
            if (L.lsum) {
                if (REAL)
                    sum = (uint64_t)__double_as_longlong(__longlong_as_double((long long)sum) +
                                                         __longlong_as_double((long long)a.p_sum[at]));
                else
                    sum += a.p_sum[at];
            }
            if (L.lmin) mn = min(mn, (uint64_t)a.p_min[at]);
            if (L.lmax) mx = max(mx, (uint64_t)a.p_max[at]);
        }
        a.p_rows[at] = rows;
        a.p_cnt[at] = cnt;
        if (L.lsum) a.p_sum[at] = sum;
        if (L.lmin) a.p_min[at] = mn;
        if (L.lmax) a.p_max[at] = mx;
    }
    const int64_t region0 = (int64_t)blockIdx.x * a.P;
    for (int q = tid; q < a.P; q += bd) a.r_count[region0 + q] = min(L.cursor[q], (uint32_t)a.cap);
    if (tid == 0) g_count_new(*gp, L.nnew[0]);
}

// One heavy row into its LDS slot l.
template <typename VT, typename VV>
__device__ __forceinline__ void p1_heavy_add(const P1Lds<VV>& L, int l, const VV& v, bool vn) {
    constexpr bool HAS_V = VTraits<VT>::has;
    constexpr bool REAL = VTraits<VT>::real;
    Part pr;
    if constexpr (HAS_V) pr = row_part<VT>(v, vn); else pr = row_part<NoValue>(NoValue{}, vn);
    if (pr.cnt) {
        atomicAdd(&L.lcnt[l], 1u);
        if (L.lsum) {
            if (REAL)
                atomicAdd((double*)&L.lsum[l], __longlong_as_double((long long)pr.sum));
            else
                atomicAdd((unsigned long long*)&L.lsum[l], (unsigned long long)pr.sum);
        }
        if (L.lmin) atomicMin((unsigned long long*)&L.lmin[l], (unsigned long long)pr.mn);
        if (L.lmax) atomicMax((unsigned long long*)&L.lmax[l], (unsigned long long)pr.mx);
    } else {
        atomicAdd(&L.lnul[l], 1u);
    }
}

// A row that cannot go to a region (NULL value, or its region is full): straight into the
// global table with atomics.
template <typename VT, typename VV>
__device__ __forceinline__ void p1_atomic_row(const GView* gp, uint32_t* nnew, uint32_t fields, uint64_t kk,
                                              uint64_t h, const VV& v, bool vn) {
    Part pr;
    if constexpr (VTraits<VT>::has) pr = row_part<VT>(v, vn); else pr = row_part<NoValue>(NoValue{}, vn);
    const GView& g = *gp;
    bool is_new = false;
    const int64_t gs = g_slot1(g, kk, false, h, &is_new);
    if (is_new) atomicAdd(nnew, 1u);
    if (gs >= 0) g_accum_atomic(g, gs, pr, fields); else g_set_lost(g);
}

// This wavefront's cold-row ring.  Head and count are the same on every lane (they only
// change by ballot counts), so the ring's control flow is uniform.
template <typename VT, typename VV>
struct P1Ring {
    uint32_t head = 0, cnt = 0;

    // Partition and write the ring's first `n` (<= 64) rows, one per lane.
    __device__ __forceinline__ void drain(const P1Lds<VV>& L, const P1Args& a, const GView* gp, uint32_t fields,
                                          uint32_t n) {
        const int lane = threadIdx.x & 63;
        const bool live = (uint32_t)lane < n;
        const uint32_t at_q = (head + (uint32_t)lane) & (kQueue - 1);
        const uint64_t kk = L.qkey[at_q];
        VV v = 0;
        if constexpr (VTraits<VT>::has) v = L.qval[at_q];
        const uint64_t h = hash64(kk);
        // part_shift is 54..64; the split shift gives 0 for 64 without a select
        const uint32_t q = (uint32_t)((h >> 1) >> (a.part_shift - 1));
        uint32_t c = ~0u;
        if (live) c = atomicAdd(&L.cursor[q], 1u);
        if (live && c < (uint32_t)a.cap) {
            put_rec<VT>(a.r_rec, ((int64_t)blockIdx.x * a.P + q) * (int64_t)a.cap + c, kk, v);
        } else if (live) {
            p1_atomic_row<VT>(gp, L.nnew, fields, kk, h, v, false);
        }
        head = (head + n) & (kQueue - 1);
    }

    // Queue the lanes' cold row (cold: this lane has one), draining 64 when that many wait.
    __device__ __forceinline__ void push(const P1Lds<VV>& L, const P1Args& a, const GView* gp, uint32_t fields,
                                         bool cold, uint64_t key, const VV& v) {
        const uint64_t m = __ballot(cold);
        if (!m) return;
        const int lane = threadIdx.x & 63;
        if (cold) {
            const uint64_t below = (1ull << lane) - 1ull;
            const uint32_t at_q = (head + cnt + (uint32_t)__popcll(m & below)) & (kQueue - 1);
            L.qkey[at_q] = key;
            if constexpr (VTraits<VT>::has) L.qval[at_q] = v;
        }
        cnt += (uint32_t)__popcll(m);
        if (cnt >= 64) {
            drain(L, a, gp, fields, 64);
            cnt -= 64;
        }
    }
};

template <typename VT, int MODE, typename KT, int FP>
__global__ void __launch_bounds__(kBlock) k_tp_phase1(const GView* gp, KeySpec ks, KeyCols kc,
                                                      const VT* __restrict__ vals,
                                                      const uint64_t* __restrict__ vnull,
                                                      const int8_t* __restrict__ mask, PredSet preds,
                                                      uint32_t vpred, int64_t n, uint32_t fields,
                                                      P1Args a) {
    extern __shared__ __align__(16) uint64_t lds[];
    constexpr bool HAS_V = VTraits<VT>::has;
    using VV = typename std::conditional<HAS_V, VT, int>::type;
    const bool f_sum = FP == kFieldsSum ? true : (fields & F_SUM) != 0;
    const bool f_min = FP == kFieldsSum ? false : (fields & F_MIN) != 0;
    const bool f_max = FP == kFieldsSum ? false : (fields & F_MAX) != 0;
    const int H = a.H, bd = blockDim.x, tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t nb = (uint32_t)(H / kBucket);
    const P1Lds<VV> L = p1_begin<VT, VV>(lds, a, f_sum, f_min, f_max);

    const int64_t tile_rows = (int64_t)bd * kP1Steps;
    const int64_t ntiles = (n + tile_rows - 1) / tile_rows;
    // rows of one wave in one step are 64 consecutive rows starting at a multiple of 64, so
    // they share one word of each null bitmap, loaded once for the wave
    const int wave_row = __builtin_amdgcn_readfirstlane(tid & ~63);
    // loads are unconditional so that a tile's loads are all in flight together: rows past
    // the end re-read the last row (and are dropped below), and a missing bitmap reads a
    // zero word
    const uint64_t* vbits = vnull ? vnull : a.zero;
    const int64_t vstep = vnull ? 1 : 0;
    const uint64_t* kbits = kc.nm[0] ? kc.nm[0] : a.zero;
    const int64_t kstep = kc.nm[0] ? 1 : 0;
    const uint64_t* cbits = preds.rnull ? preds.rnull : a.zero;
    const int64_t cstep = preds.rnull ? 1 : 0;
    struct Tile {
        uint64_t key[kP1Steps], vword[kP1Steps], kword[kP1Steps], cword[kP1Steps];
        int64_t cv[kP1Steps];
        bool kn[kP1Steps];
        VV val[kP1Steps];
    };
    auto load_tile = [&](int64_t tile, Tile& t) {
#pragma unroll
        for (int k = 0; k < kP1Steps; k++) {
            const int64_t r0 = min(tile * tile_rows + (int64_t)k * bd + wave_row, n - 1);
            const int64_t r = min(r0 + lane, n - 1);
            t.vword[k] = vbits[(r0 >> 6) * vstep];
            t.kword[k] = 0;
            t.kn[k] = false;
            if constexpr (kTypedKey<KT>) {
                t.kword[k] = kbits[(r0 >> 6) * kstep];
                // streamed once: non-temporal, so that the lines of the cold regions being
                // filled stay in the L2 instead of being pushed out by the column reads
                t.key[k] = (uint64_t)(int64_t)__builtin_nontemporal_load((const KT*)kc.col[0] + r);
            } else if constexpr (kPairKey<KT>) {
                // the key spec's packing of two int32 fields at bits 0 and 32 (no null bits)
                const uint32_t lo = (uint32_t)__builtin_nontemporal_load((const int32_t*)kc.col[0] + r);
                const uint32_t hi = (uint32_t)__builtin_nontemporal_load((const int32_t*)kc.col[1] + r);
                t.key[k] = (uint64_t)lo | ((uint64_t)hi << 32);
            } else {
                uint64_t kk[1];
                load_key<1>(ks, kc, r, kk, t.kn[k]);
                t.key[k] = kk[0];
            }
            if constexpr (HAS_V) t.val[k] = __builtin_nontemporal_load(vals + r); else t.val[k] = 0;
            t.cv[k] = 0;
            t.cword[k] = 0;
            if (MODE == 4) {
                t.cv[k] = range_col(preds, r);
                t.cword[k] = cbits[(r0 >> 6) * cstep];
            }
        }
    };
    // A tile's rows after the WHERE: key (0 for a NULL key), value, and per step whether the
    // row passes and whether its key / value is NULL.
    struct Rows {
        uint64_t key[kP1Steps];
        VV val[kP1Steps];
        uint32_t pass, vnull, knull;
    };
    auto filter_tile = [&](int64_t tile, const Tile& t, Rows& o) {
        o.pass = o.vnull = o.knull = 0;
#pragma unroll
        for (int k = 0; k < kP1Steps; k++) {
            const int64_t r = tile * tile_rows + (int64_t)k * bd + tid;
            const bool vn = (t.vword[k] >> lane) & 1;
            bool kn = t.kn[k];
            uint64_t kk = t.key[k];
            if constexpr (kTypedKey<KT>) {
                kn = (t.kword[k] >> lane) & 1;
                if (kn) kk = 0;
            }
            o.key[k] = kk;
            o.val[k] = t.val[k];
            const bool pass = (r < n) & row_passes<VT, MODE>(preds, vpred, mask, r, t.val[k], vn, t.cv[k],
                                                             (t.cword[k] >> lane) & 1);
            o.pass |= (uint32_t)pass << k;
            o.vnull |= (uint32_t)vn << k;
            o.knull |= (uint32_t)kn << k;
        }
    };
    P1Ring<VT, VV> ring;
    auto aggregate_rows = [&](const Rows& o) {
        // every step's bucket is read first (a NULL key reads key 0's; the NULL key and the
        // key equal to the empty marker then take their side slots by selection)
        int ls[kP1Steps];
#pragma unroll
        for (int k = 0; k < kP1Steps; k++) {
            const uint64_t kk = o.key[k];
            int l = heavy_find(L.lkey, nb, kk);
            l = kk == kEmpty ? H + 1 : l;
            ls[k] = ((o.knull >> k) & 1) ? H : l;
        }
#pragma unroll
        for (int k = 0; k < kP1Steps; k++) {
            const bool pass = (o.pass >> k) & 1, vn = (o.vnull >> k) & 1;
            if (pass && ls[k] >= 0) p1_heavy_add<VT>(L, ls[k], o.val[k], vn);
        }
        // cold rows with a value go to the ring; a NULL value takes the atomic path
#pragma unroll
        for (int k = 0; k < kP1Steps; k++) {
            const bool pass = (o.pass >> k) & 1, vn = (o.vnull >> k) & 1;
            const bool cold = pass & (ls[k] < 0);
            if (cold && vn) p1_atomic_row<VT>(gp, L.nnew, fields, o.key[k], hash64(o.key[k]), o.val[k], true);
            ring.push(L, a, gp, fields, cold && !vn, o.key[k], o.val[k]);
        }
    };
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        Tile t;
        Rows o;
        load_tile(tile, t);
        filter_tile(tile, t, o);
        aggregate_rows(o);
    }
    if (ring.cnt) ring.drain(L, a, gp, fields, ring.cnt);
    p1_end<VTraits<VT>::real>(L, a, gp);
}

// ---- phase 1, streaming form ----------------------------------------------------------------
//
// The same work as k_tp_phase1 for one typed key column (int32 or int64, nullable) or two
// non-NULL int32 key columns packed into one word, with no WHERE or a WHERE folded into an
// interval (on the value column or on one other integer column).  A wavefront takes chunks of
// kSChunk rows in parts of kSPart rows; lane l holds rows 2l and 2l+1 of every part, so every
// load instruction reads one contiguous 512 B (4-byte column) or 1 KB (8-byte column) range
// with 8- or 16-byte lanes, and a null bitmap costs one 32-bit load per part.  The loads are
// buffer loads bounded by the resource's size: rows past the end read as zero and an absent
// bitmap is a resource of size zero, so no load needs a bounds
// test or a clamp, and the next chunk's loads are issued before the current chunk is
// aggregated.  Offsets are 32-bit bytes: every column of the block stays below 2^31 bytes.

constexpr int kSRows = 4;                 // rows per lane per chunk
constexpr int kSParts = kSRows / 2;        // parts of a chunk: lane l holds rows 2l, 2l+1 of each
constexpr int kSChunk = 64 * kSRows;      // rows per wavefront per chunk
constexpr int kSPart = 128;               // rows per part

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, int64_t bytes) {
    // word 3 = 0x00020000: the untyped 32-bit data format of raw buffer accesses on gfx9
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// The 32-bit words of one lane's kSRows rows of a column of E-byte elements: row j of the lane
// in w[j] (E = 4) or w[2j], w[2j+1] (E = 8).
template <int E>
struct LaneWords {
    uint32_t w[E * kSRows / 4];
};

// Every part of a lane's rows, the first at byte offset off, the next kSPart * E bytes further
// each (non-temporal: the columns are streamed once, and the lines of the cold regions being
// filled should stay in the L2).
__device__ __forceinline__ void ld_lane(__amdgpu_buffer_rsrc_t r, uint32_t off, LaneWords<4>& t) {
#pragma unroll
    for (int q = 0; q < kSParts; q++) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + q * kSPart * 4), 0, 2);
        t.w[2 * q] = v[0];
        t.w[2 * q + 1] = v[1];
    }
}
__device__ __forceinline__ void ld_lane(__amdgpu_buffer_rsrc_t r, uint32_t off, LaneWords<8>& t) {
#pragma unroll
    for (int q = 0; q < kSParts; q++) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(off + q * kSPart * 8), 0, 2);
        t.w[4 * q] = v[0];
        t.w[4 * q + 1] = v[1];
        t.w[4 * q + 2] = v[2];
        t.w[4 * q + 3] = v[3];
    }
}

template <typename T>
__device__ __forceinline__ T word_elem(const uint32_t* w, int j) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
        return __builtin_bit_cast(T, u);
    } else {
        return __builtin_bit_cast(T, w[j]);
    }
}

// A null bitmap's words for every part of a lane's rows: the lane's two rows of a part lie in
// one 32-bit word (2l is even), each part's word kSPart / 32 words after the previous one's.
struct LaneBits {
    uint32_t w[kSParts];
    __device__ __forceinline__ bool bit(int j, uint32_t sh) const {
        return (w[j >> 1] >> (sh + (uint32_t)(j & 1))) & 1u;
    }
};
__device__ __forceinline__ LaneBits no_bits() {
    LaneBits b;
#pragma unroll
    for (int q = 0; q < kSParts; q++) b.w[q] = 0u;
    return b;
}
__device__ __forceinline__ LaneBits ld_bits(__amdgpu_buffer_rsrc_t r, uint32_t boff) {
    LaneBits b;
#pragma unroll
    for (int q = 0; q < kSParts; q++)
        b.w[q] = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(boff + q * (kSPart / 8)), 0, 0);
    return b;
}

template <typename VT, int MODE, typename KT, int FP>
__global__ void __launch_bounds__(kBlock) k_tp_phase1s(const GView* gp, KeyCols kc, const VT* __restrict__ vals,
                                                       const uint64_t* __restrict__ vnull, PredSet preds,
                                                       int32_t n, uint32_t fields, P1Args a) {
    extern __shared__ __align__(16) uint64_t lds[];
    constexpr bool HAS_V = VTraits<VT>::has;
    constexpr bool PAIR = kPairKey<KT>;
    using VV = typename std::conditional<HAS_V, VT, int>::type;
    constexpr int KE = PAIR ? 4 : (int)sizeof(KT);          // bytes per row of a key column
    constexpr int VE = HAS_V ? (int)sizeof(VV) : 4;         // bytes per row of the value column
    const bool f_sum = FP == kFieldsSum ? true : (fields & F_SUM) != 0;
    const bool f_min = FP == kFieldsSum ? false : (fields & F_MIN) != 0;
    const bool f_max = FP == kFieldsSum ? false : (fields & F_MAX) != 0;
    const int H = a.H, tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t nb = (uint32_t)(H / kBucket);
    const P1Lds<VV> L = p1_begin<VT, VV>(lds, a, f_sum, f_min, f_max);
    const int32_t K = __builtin_amdgcn_readfirstlane(min(a.meta[0], a.kmax));

    const int64_t bitmap_bytes = ((int64_t)n + 63) / 64 * 8;
    const __amdgpu_buffer_rsrc_t rk0 = make_rsrc(kc.col[0], (int64_t)n * KE);
    const __amdgpu_buffer_rsrc_t rk1 = make_rsrc(PAIR ? kc.col[1] : kc.col[0], PAIR ? (int64_t)n * KE : 0);
    const __amdgpu_buffer_rsrc_t rkn = make_rsrc(kc.nm[0], (!PAIR && kc.nm[0]) ? bitmap_bytes : 0);
    const __amdgpu_buffer_rsrc_t rv = make_rsrc(vals, HAS_V ? (int64_t)n * VE : 0);
    const __amdgpu_buffer_rsrc_t rvn = make_rsrc(vnull, vnull ? bitmap_bytes : 0);
    const int CE = preds.rwide ? 8 : 4;
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(preds.rcol, MODE == 4 ? (int64_t)n * CE : 0);
    const __amdgpu_buffer_rsrc_t rcn = make_rsrc(preds.rnull, (MODE == 4 && preds.rnull) ? bitmap_bytes : 0);

    struct Chunk {
        LaneWords<KE> k0, k1;
        LaneWords<VE> v;
        LaneWords<8> c;
        LaneBits vb, kb, cb;
    };
    // the chunk's loads; a chunk past the end reads zeros
    auto load = [&](uint32_t chunk, Chunk& t) {
        const uint32_t row0 = chunk * (uint32_t)kSChunk + (uint32_t)lane * 2;
        ld_lane(rk0, row0 * KE, t.k0);
        if constexpr (PAIR) ld_lane(rk1, row0 * KE, t.k1);
        if constexpr (HAS_V) ld_lane(rv, row0 * VE, t.v);
        const uint32_t boff = (row0 >> 5) * 4;
        t.vb = ld_bits(rvn, boff);
        t.kb = PAIR ? no_bits() : ld_bits(rkn, boff);
        t.cb = no_bits();
        if constexpr (MODE == 4) {
            if (preds.rwide) {
                ld_lane(rc, row0 * 8, t.c);
            } else {
                LaneWords<4> c4;
                ld_lane(rc, row0 * 4, c4);
#pragma unroll
                for (int j = 0; j < kSRows; j++) t.c.w[j] = c4.w[j];
            }
            t.cb = ld_bits(rcn, boff);
        }
    };
    P1Ring<VT, VV> ring;
    auto process = [&](uint32_t chunk, const Chunk& t) {
        const uint32_t row0 = chunk * (uint32_t)kSChunk + (uint32_t)lane * 2;
        const uint32_t sh = row0 & 31;
        uint64_t key[kSRows];
        VV val[kSRows];
        int ls[kSRows];
        uint32_t pass = 0, vnl = 0, knl = 0;
#pragma unroll
        for (int j = 0; j < kSRows; j++) {
            const bool vn = t.vb.bit(j, sh);
            const bool kn = t.kb.bit(j, sh);
            uint64_t kk;
            if constexpr (PAIR)
                kk = (uint64_t)t.k0.w[j] | ((uint64_t)t.k1.w[j] << 32);
            else
                kk = (uint64_t)(int64_t)word_elem<KT>(t.k0.w, j);
            kk = kn ? 0ull : kk;
            VV v = 0;
            if constexpr (HAS_V) v = word_elem<VV>(t.v.w, j);
            int64_t cv = 0;
            if constexpr (MODE == 4)
                cv = preds.rwide ? word_elem<int64_t>(t.c.w, j) : (int64_t)word_elem<int32_t>(t.c.w, j);
            const bool cn = t.cb.bit(j, sh);
            const bool in_block = row0 + (uint32_t)((j >> 1) * kSPart + (j & 1)) < (uint32_t)n;
            const bool ok = in_block & row_passes<VT, MODE>(preds, 0u, nullptr, 0, v, vn, cv, cn);
            key[j] = kk;
            val[j] = v;
            pass |= (uint32_t)ok << j;
            vnl |= (uint32_t)vn << j;
            knl |= (uint32_t)kn << j;
        }
        // every row's bucket is read first (a NULL key reads key 0's; the NULL key and the key
        // equal to the empty marker then take their side slots by selection); with no heavy key
        // the table is not read
#pragma unroll
        for (int j = 0; j < kSRows; j++) {
            int l = K ? heavy_find(L.lkey, nb, key[j]) : -1;
            l = key[j] == kEmpty ? H + 1 : l;
            ls[j] = ((knl >> j) & 1) ? H : l;
        }
#pragma unroll
        for (int j = 0; j < kSRows; j++)
            if (((pass >> j) & 1) && ls[j] >= 0) p1_heavy_add<VT>(L, ls[j], val[j], (vnl >> j) & 1);
        // cold rows with a value go to the ring; a NULL value takes the atomic path
#pragma unroll
        for (int j = 0; j < kSRows; j++) {
            const bool vn = (vnl >> j) & 1;
            const bool cold = ((pass >> j) & 1) & (ls[j] < 0);
            if (cold && vn) p1_atomic_row<VT>(gp, L.nnew, fields, key[j], hash64(key[j]), val[j], true);
            ring.push(L, a, gp, fields, cold && !vn, key[j], val[j]);
        }
    };
    // chunks are dealt to the wavefronts of the grid in turn; each wavefront loads its next
    // chunk before it aggregates the current one (two register sets, used alternately)
    const uint32_t nchunks = (uint32_t)(((int64_t)n + kSChunk - 1) / kSChunk);
    const uint32_t stride = gridDim.x * (uint32_t)(kBlock / 64);
    uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (tid >> 6)));
    Chunk A, B;
    load(c, A);
    while (c < nchunks) {
        load(c + stride, B);
        process(c, A);
        c += stride;
        if (c >= nchunks) break;
        load(c + stride, A);
        process(c, B);
        c += stride;
    }
    if (ring.cnt) ring.drain(L, a, gp, fields, ring.cnt);
    p1_end<VTraits<VT>::real>(L, a, gp);
}

// ---- phase 2 -------------------------------------------------------------------------------

struct P2Args {
    const uint64_t* heavy;
    const int32_t* meta;
    int32_t kmax, H2, P, cap, W;
    const uint32_t* p_cnt;
    const uint32_t* p_rows;
    const uint64_t* p_sum;
    const uint64_t* p_min;
    const uint64_t* p_max;
    const uint32_t* r_rec;
    const uint32_t* r_count;
    int32_t split;  // log2(sub-tables per partition), owned phase 2 only
};

__host__ __device__ inline size_t p2_lds_bytes(int32_t H2, int32_t W, uint32_t fields) {
    const size_t T = (size_t)H2 + 2;
    size_t b = (size_t)H2 * 8;
    b += T * 8 * (((fields & F_SUM) ? 1 : 0) + ((fields & F_MIN) ? 1 : 0) + ((fields & F_MAX) ? 1 : 0));
    b += T * 4 * 2;
    b += 16;
    (void)W;
    return b;
}

// Home slot of a key in a partition's LDS table.  The partition already fixes the top bits of
// the key's 64-bit hash, so the table uses the folded multiplicative hash instead (as the
// phase-1 buckets do); the 64-bit hash is computed once per key, when it is merged.
__device__ __forceinline__ uint32_t p2_home(uint64_t k, int32_t H2) {
    const uint32_t f = ((uint32_t)k ^ (uint32_t)(k >> 32)) * 0x9e3779b1u;
    return __umulhi(f, (uint32_t)H2);
}

template <typename VT, int FP>
__global__ void __launch_bounds__(kBlock) k_tp_phase2(GView g, uint32_t fields, P2Args a) {
    extern __shared__ __align__(16) uint64_t lds[];
    constexpr bool HAS_V = VTraits<VT>::has;
    constexpr bool REAL = VTraits<VT>::real;
    const int H = a.H2, T = H + 2, bd = blockDim.x, tid = threadIdx.x, part = blockIdx.x;
    uint64_t* tkey = lds;
    uint64_t* p = tkey + H;
    const bool f_sum = FP == kFieldsSum ? true : (fields & F_SUM) != 0;
    const bool f_min = FP == kFieldsSum ? false : (fields & F_MIN) != 0;
    const bool f_max = FP == kFieldsSum ? false : (fields & F_MAX) != 0;
    uint64_t* tsum = f_sum ? p : nullptr;
    p += f_sum ? T : 0;
    uint64_t* tmin = f_min ? p : nullptr;
    p += f_min ? T : 0;
    uint64_t* tmax = f_max ? p : nullptr;
    p += f_max ? T : 0;
    uint32_t* tcnt = (uint32_t*)p;    // non-NULL values
    uint32_t* trows = tcnt + T;       // rows beyond tcnt (rows with a NULL value)
    uint32_t* misc = trows + T;       // [0] used slots, [1] new keys, [2] overflow rows, [3] rows

    auto clear_slot = [&](int s) {
        if (s < H) tkey[s] = kEmpty;
        tcnt[s] = 0;
        trows[s] = 0;
        if (tsum) tsum[s] = 0;
        if (tmin) tmin[s] = ~0ull;
        if (tmax) tmax[s] = 0;
    };
    for (int s = tid; s < T; s += bd) clear_slot(s);
    if (tid == 0) {
        misc[0] = 0;
        misc[1] = 0;
        misc[2] = 0;
        misc[3] = 0;
    }
    __syncthreads();

    // A new key takes a slot only while fewer than `limit` slots are in use; past that the
    // caller aggregates the row straight into the global table.  The count is read before the
    // claim and raised after it, so up to one claim per thread may pass the check together;
    // limit <= H - blockDim keeps empty slots in the table then (and a probe gives up after H
    // slots in any case).  Linear probing without deletion: once an empty slot is seen, the key
    // is not further along the chain.
    const uint32_t limit = (uint32_t)max(H / 2, min(H * 4 / 5, H - bd));
    auto insert_from = [&](uint64_t key, uint32_t s) -> int {
        for (int it = 0; it < H; it++) {
            const uint64_t cur = tkey[s];
            if (cur == key) return (int)s;
            if (cur == kEmpty) {
                if (__hip_atomic_load(&misc[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= limit) return -1;
                const unsigned long long old = atomicCAS((unsigned long long*)&tkey[s],
                                                         (unsigned long long)kEmpty, (unsigned long long)key);
                if (old == kEmpty) {
                    atomicAdd(&misc[0], 1u);
                    return (int)s;
                }
                if (old == key) return (int)s;
            }
            s = (s + 1 == (uint32_t)H) ? 0u : s + 1;
        }
        return -1;
    };
    // homes lie in the first H - kWindow slots, so the kWindow slots from a home never wrap
    constexpr int kWindow = 4;
    const int Hh = H - 8;
    auto insert = [&](uint64_t key) -> int { return insert_from(key, p2_home(key, Hh)); };
    // rows of keys that found no LDS slot
    auto add_global = [&](uint64_t key, const Part& pr) {
        bool is_new = false;
        const int64_t gs = g_slot1(g, key, false, hash64(key), &is_new);
        if (is_new) atomicAdd(&misc[1], 1u);
        if (gs >= 0) g_accum_atomic(g, gs, pr, fields); else g_set_lost(g);
        atomicAdd(&misc[2], 1u);
    };
    auto add = [&](int s, uint32_t rows, uint32_t cnt, uint64_t sum, uint64_t mn, uint64_t mx) {
        if (rows) atomicAdd(&trows[s], rows);
        if (cnt) {
            atomicAdd(&tcnt[s], cnt);
            if (tsum) {
                if (REAL)
                    atomicAdd((double*)&tsum[s], __longlong_as_double((long long)sum));
                else
                    atomicAdd((unsigned long long*)&tsum[s], (unsigned long long)sum);
            }
            if (tmin) atomicMin((unsigned long long*)&tmin[s], (unsigned long long)mn);
            if (tmax) atomicMax((unsigned long long*)&tmax[s], (unsigned long long)mx);
        }
    };
    auto part_of_slot = [&](int s) {
        Part q;
        q.cnt = tcnt[s];
        q.cntall = trows[s] + tcnt[s];
        q.sum = tsum ? tsum[s] : 0;
        q.mn = tmin ? tmin[s] : ~0ull;
        q.mx = tmax ? tmax[s] : 0;
        return q;
    };
    // fold the LDS table into this partition's sub-table (this workgroup is its only writer)
    auto merge = [&]() {
        // rows that went straight to the global table (atomics at device scope) are complete
        // and visible to this workgroup's plain reads below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        for (int s = tid; s < H; s += bd) {
            const uint64_t key = tkey[s];
            if (key == kEmpty) continue;
            bool is_new = false;
            const int64_t gs = g_slot1_claim(g, key, hash64(key), &is_new);
            if (is_new) atomicAdd(&misc[1], 1u);
            if (gs < 0) g_set_lost(g);
            else if (is_new) g_store_new(g, gs, part_of_slot(s), fields);
            else g_accum_owned(g, gs, part_of_slot(s), fields);
            clear_slot(s);
        }
        __syncthreads();
        if (tid == 0) misc[0] = 0;
        __syncthreads();
    };

    // heavy partials of this partition's heavy keys, summed over the phase-1 workgroups: a
    // thread owns one key and a strided subset of the workgroups, so neighbouring lanes read
    // neighbouring keys of one partial row, and each thread adds one combined partial
    const int32_t hk0 = a.meta[2 + part], hk1 = a.meta[2 + part + 1];
    const int32_t nk = hk1 - hk0;
    const int32_t stride_k = a.kmax + 2;
    if (nk > 0) {
        const int per = nk <= bd ? bd / nk : 1;  // threads per key
        const int span = nk <= bd ? per * nk : bd;
        if (tid < span) {
            const int w0 = nk <= bd ? tid / nk : 0;
            for (int ki = nk <= bd ? tid % nk : tid; ki < nk; ki += (nk <= bd ? nk : bd)) {
                uint32_t rows = 0, cnt = 0;
                uint64_t isum = 0, mn = ~0ull, mx = 0;
                double dsum = 0.0;
#pragma unroll 8
                for (int w = w0; w < a.W; w += per) {
                    const int64_t at = (int64_t)w * stride_k + hk0 + ki;
                    rows += a.p_rows[at];
                    cnt += a.p_cnt[at];
                    if (tsum) {
                        if (REAL) dsum += __longlong_as_double((long long)a.p_sum[at]);
                        else isum += a.p_sum[at];
                    }
                    if (tmin) mn = min(mn, (uint64_t)a.p_min[at]);
                    if (tmax) mx = max(mx, (uint64_t)a.p_max[at]);
                }
                if (!rows) continue;
                const uint64_t hsum = REAL ? (uint64_t)__double_as_longlong(dsum) : isum;
                const int s = insert(a.heavy[hk0 + ki]);
                if (s >= 0) {
                    add(s, rows - cnt, cnt, hsum, mn, mx);
                } else {
                    Part q;
                    q.cnt = cnt;
                    q.cntall = rows;
                    q.sum = hsum;
                    q.mn = mn;
                    q.mx = mx;
                    add_global(a.heavy[hk0 + ki], q);
                }
            }
        }
    }
    if (part == 0) {  // NULL key and key == kEmpty: side slots
        for (int64_t j = tid; j < 2 * (int64_t)a.W; j += bd) {
            const int32_t side = (int32_t)(j / a.W), w = (int32_t)(j % a.W);
            const int64_t at = (int64_t)w * stride_k + a.kmax + side;
            const uint32_t rows = a.p_rows[at];
            if (!rows) continue;
            add(H + side, rows - a.p_cnt[at], a.p_cnt[at], tsum ? a.p_sum[at] : 0,
                tmin ? a.p_min[at] : ~0ull, tmax ? a.p_max[at] : 0);
        }
    }

    // Cold rows of this partition: in each round every wave takes the partition's regions of
    // kRegionsPerWave phase-1 workgroups and walks them as one run of rows, 64 x kP2Rows rows
    // at a time, loading them all before inserting any.  A round in which many rows found no
    // LDS slot means the partition holds more keys than the table: the table is then merged
    // and cleared before the next round.
    constexpr int kRegionsPerWave = 4;
    const int lane = tid & 63, wave = tid >> 6, nwaves = bd >> 6;
    for (int r0 = 0; r0 < a.W; r0 += nwaves * kRegionsPerWave) {
        __syncthreads();
        // read into scalar registers: a barrier-containing branch on an LDS value must be
        // visibly uniform, or the compiler may split the wave around it
        const bool flush = __builtin_amdgcn_readfirstlane((int)(misc[2] > 0 && (uint64_t)misc[2] * 8 > misc[3])) != 0;
        __syncthreads();
        if (flush) merge();
        if (tid == 0) {
            misc[2] = 0;
            misc[3] = 0;
        }
        __syncthreads();
        // region g of this wave starts at row pre[g] of the run
        int64_t rbase[kRegionsPerWave];
        int32_t pre[kRegionsPerWave + 1];
        pre[0] = 0;
#pragma unroll
        for (int g = 0; g < kRegionsPerWave; g++) {
            const int w = r0 + wave * kRegionsPerWave + g;
            const int32_t c = w < a.W ? (int32_t)a.r_count[(int64_t)w * a.P + part] : 0;
            rbase[g] = ((int64_t)w * a.P + part) * (int64_t)a.cap;
            pre[g + 1] = pre[g] + c;
        }
        const int32_t cnt = pre[kRegionsPerWave];
        if (lane == 0 && cnt) atomicAdd(&misc[3], (uint32_t)cnt);
        for (int32_t j0 = 0; j0 < cnt; j0 += 64 * kP2Rows) {
            uint64_t key[kP2Rows];
            typename std::conditional<HAS_V, VT, int>::type val[kP2Rows];
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                const int32_t j = min(j0 + i * 64 + lane, cnt - 1);
                int64_t at = rbase[0] + j;
#pragma unroll
                for (int g = 1; g < kRegionsPerWave; g++)
                    if (j >= pre[g]) at = rbase[g] + (j - pre[g]);
                get_rec<VT>(a.r_rec, at, key[i], val[i]);
            }
            // the kWindow slots from the home of every row at once: a key present in the table
            // sits there unless its chain is longer (no deletions, so any match is the key's)
            uint32_t home[kP2Rows];
            int hit[kP2Rows];
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) home[i] = p2_home(key[i], Hh);
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                int h = -1;
#pragma unroll
                for (int w = kWindow - 1; w >= 0; w--) h = tkey[home[i] + w] == key[i] ? (int)home[i] + w : h;
                hit[i] = h;
            }
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                if (j0 + i * 64 + lane >= cnt) continue;
                Part pr;
                if constexpr (HAS_V) {
                    pr = row_part<VT>(val[i], false);
                } else {
                    pr = row_part<NoValue>(NoValue{}, false);
                }
                int s = hit[i];  // only keys further along, or new, branch
                if (s < 0) s = insert_from(key[i], home[i]);
                if (s >= 0)
                    add(s, 0, 1, pr.sum, pr.mn, pr.mx);
                else
                    add_global(key[i], pr);
            }
        }
    }
    merge();
    if (part == 0 && tid < 2) {
        const int s = H + tid;
        if (trows[s] + tcnt[s]) {  // rows reached the side slot
            const uint64_t gs = g.cap + tid;
            if (g_claim_side(g, gs) && tid == 1) atomicAdd(&misc[1], 1u);
            g_accum_owned(g, (int64_t)gs, part_of_slot(s), fields);
        }
    }
    __syncthreads();
    if (tid == 0) g_count_new(g, misc[1]);
}

// ---- phase 2, owned sub-table -------------------------------------------------------------

// Phase 2 when a partition's sub-table fits the LDS.  The workgroup owns its sub-table: it
// loads the sub-table's keys, folds the heavy partials and the partition's cold rows into
// deltas at the global table's own slot positions (the same home slot and probe order as
// g_slot1), and adds the deltas to the sub-table's states.  Every global access is coalesced
// and no row needs a global probe or atomic.  Counts are 32-bit deltas (one flush holds at
// most 2^30 rows); the global states stay 64-bit.  The kernel has no static LDS, so the
// sub-table may use all of it.
template <typename VT, int FP>
__global__ void __launch_bounds__(kBlock) k_tp_phase2_owned(GView g, uint32_t fields, P2Args a) {
    extern __shared__ __align__(16) uint64_t lds[];
    constexpr bool HAS_V = VTraits<VT>::has;
    constexpr bool REAL = VTraits<VT>::real;
    const bool f_sum = FP == kFieldsSum ? true : (fields & F_SUM) != 0;
    const bool f_min = FP == kFieldsSum ? false : (fields & F_MIN) != 0;
    const bool f_max = FP == kFieldsSum ? false : (fields & F_MAX) != 0;
    const bool f_all = (fields & F_CNTALL) != 0;
    const int tid = threadIdx.x, bd = blockDim.x;
    // this workgroup's sub-table, and the partition whose rows hold its keys.  With several
    // sub-tables per partition (split > 0), the sub-tables of one partition go to workgroups
    // of one die (dispatch deals workgroups to the 8 dies in turn), so the partition's rows
    // are read from that die's L2 after the first of them.
    int part = blockIdx.x;
    if (a.split) {
        const int die = blockIdx.x & 7, i = blockIdx.x >> 3;
        part = ((((i >> a.split) << 3) + die) << a.split) | (i & ((1 << a.split) - 1));
    }
    const int rpart = part >> a.split;
    // sub-table of a key hash (a two-step shift, so 64 bits of shift give 0)
    auto sub_of = [&](uint64_t h) { return (int)((h >> 1) >> (g.part_shift - 1)); };
    const uint32_t S = (uint32_t)g.sub_mask + 1;
    const int64_t base = (int64_t)part << g.sub_bits;
    // 8-byte arrays first, then the 4-byte ones
    uint64_t* okey = lds;
    unsigned long long* p = (unsigned long long*)(okey + S);
    unsigned long long* osum = f_sum ? p : nullptr;
    p += f_sum ? S : 0;
    unsigned long long* omin = f_min ? p : nullptr;
    p += f_min ? S : 0;
    unsigned long long* omax = f_max ? p : nullptr;
    p += f_max ? S : 0;
    uint32_t* ocnt = (uint32_t*)p;
    uint32_t* oall = f_all ? ocnt + S : nullptr;

    // kFill slots per thread at a time, their loads issued together
    constexpr int kFill = 4;
    for (uint32_t s0 = tid; s0 < S; s0 += kFill * bd) {
        uint64_t k[kFill];
#pragma unroll
        for (int u = 0; u < kFill; u++) {
            const uint32_t s = s0 + u * bd;
            k[u] = s < S ? g.keys[base + s] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kFill; u++) {
            const uint32_t s = s0 + u * bd;
            if (s >= S) continue;
            okey[s] = k[u];
            ocnt[s] = 0;
            if (oall) oall[s] = 0;
            if (osum) osum[s] = 0;
            if (omin) omin[s] = ~0ull;
            if (omax) omax[s] = 0;
        }
    }
    __syncthreads();

    uint32_t nnew = 0;  // keys this thread added to the table
    // slot of a key (claimed when absent), probing from h & sub_mask as g_slot1 does
    // (free: slot s was seen empty, so the claim starts with the CAS)
    auto claim_from = [&](uint64_t key, uint32_t s, bool free = false) -> int {
        for (uint32_t it = 0; it < S; it++) {
            const uint64_t cur = free ? kEmpty : okey[s];
            free = false;
            if (cur == key) return (int)s;
            if (cur == kEmpty) {
                const unsigned long long old = atomicCAS((unsigned long long*)&okey[s], (unsigned long long)kEmpty,
                                                         (unsigned long long)key);
                if (old == kEmpty) {
                    nnew++;
                    return (int)s;
                }
                if (old == key) return (int)s;
            }
            s = (s + 1) & (S - 1);
        }
        return -1;
    };
    auto add = [&](int s, uint32_t rows, uint32_t cnt, uint64_t sum, uint64_t mn, uint64_t mx) {
        if (oall && rows) atomicAdd(&oall[s], rows);
        if (cnt) {
            atomicAdd(&ocnt[s], cnt);
            if (osum) {
                if (REAL)
                    atomicAdd((double*)&osum[s], __longlong_as_double((long long)sum));
                else
                    atomicAdd(&osum[s], (unsigned long long)sum);
            }
            if (omin) atomicMin(&omin[s], (unsigned long long)mn);
            if (omax) atomicMax(&omax[s], (unsigned long long)mx);
        }
    };

    // heavy partials of this partition's heavy keys, summed over the phase-1 workgroups (a
    // thread owns one key and a strided subset of the workgroups)
    const int32_t hk0 = a.meta[2 + rpart], hk1 = a.meta[2 + rpart + 1];
    const int32_t nk = hk1 - hk0;
    const int32_t stride_k = a.kmax + 2;
    if (nk > 0) {
        const int per = nk <= bd ? bd / nk : 1;
        const int span = nk <= bd ? per * nk : bd;
        if (tid < span) {
            const int w0 = nk <= bd ? tid / nk : 0;
            for (int ki = nk <= bd ? tid % nk : tid; ki < nk; ki += (nk <= bd ? nk : bd)) {
                const uint64_t key = a.heavy[hk0 + ki];
                const uint64_t kh = hash64(key);
                if (sub_of(kh) != part) continue;
                uint64_t rows = 0, cnt = 0, isum = 0, mn = ~0ull, mx = 0;
                double dsum = 0.0;
#pragma unroll 8
                for (int w = w0; w < a.W; w += per) {
                    const int64_t at = (int64_t)w * stride_k + hk0 + ki;
                    rows += a.p_rows[at];
                    cnt += a.p_cnt[at];
                    if (osum) {
                        if (REAL) dsum += __longlong_as_double((long long)a.p_sum[at]);
                        else isum += a.p_sum[at];
                    }
                    if (omin) mn = min(mn, (uint64_t)a.p_min[at]);
                    if (omax) mx = max(mx, (uint64_t)a.p_max[at]);
                }
                if (!rows) continue;
                const int s = claim_from(key, (uint32_t)(kh & g.sub_mask));
                if (s >= 0) add(s, (uint32_t)rows, (uint32_t)cnt, REAL ? (uint64_t)__double_as_longlong(dsum) : isum, mn, mx);
                else g_set_lost(g);
            }
        }
    }

    // cold rows: every wave walks the partition's regions of kRegionsPerWave phase-1
    // workgroups at a time as one run of rows.  The wave index is read into a scalar register,
    // so the run's counts and offsets are scalar; the next run's counts are loaded while this
    // run is walked.
    constexpr int kRegionsPerWave = 4;
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwaves = bd >> 6;
    const int32_t stride = a.P * a.cap;  // between the regions of one partition (host-checked)
    int32_t cn[kRegionsPerWave];
    auto counts = [&](int r0) {
#pragma unroll
        for (int q = 0; q < kRegionsPerWave; q++) {
            const int w = r0 + q;
            cn[q] = w < a.W ? (int32_t)a.r_count[(int64_t)w * a.P + rpart] : 0;
        }
    };
    counts(wave * kRegionsPerWave);
    for (int r0 = wave * kRegionsPerWave; r0 < a.W; r0 += nwaves * kRegionsPerWave) {
        // row j of the run lies in region q = #{q' >= 1 : j >= pre[q']}, at record
        // rbase + j + the jumps of those q' (a jump skips the unused rest of a region)
        const int64_t rbase = ((int64_t)r0 * a.P + rpart) * (int64_t)a.cap;
        int32_t pre[kRegionsPerWave], jump[kRegionsPerWave];
        pre[0] = 0;
        jump[0] = 0;
#pragma unroll
        for (int q = 1; q < kRegionsPerWave; q++) {
            pre[q] = pre[q - 1] + cn[q - 1];
            jump[q] = stride - cn[q - 1];
        }
        const int32_t cnt = pre[kRegionsPerWave - 1] + cn[kRegionsPerWave - 1];
        counts(r0 + nwaves * kRegionsPerWave);
        for (int32_t j0 = 0; j0 < cnt; j0 += 64 * kP2Rows) {
            uint64_t key[kP2Rows];
            typename std::conditional<HAS_V, VT, int>::type val[kP2Rows];
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                const int32_t j = min(j0 + i * 64 + lane, cnt - 1);
                int32_t off = j;
#pragma unroll
                for (int q = 1; q < kRegionsPerWave; q++) off += j >= pre[q] ? jump[q] : 0;
                const int64_t at = rbase + off;
                get_rec<VT>(a.r_rec, at, key[i], val[i]);
            }
            // a window of four slots around every row's home, read as two 128-bit words from
            // the even slot at or below it: a key present in the table sits there unless its
            // chain is longer (no deletions, so any match is the key's slot).  Home slot, with
            // bit 31 set for a row of another sub-table of the partition.
            uint32_t home[kP2Rows], from[kP2Rows];
            int hit[kP2Rows];
            uint32_t free_at = 0;  // bit i: from[i] was seen empty
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                const uint64_t h = hash64(key[i]);
                home[i] = (uint32_t)(h & g.sub_mask) | ((uint32_t)(sub_of(h) != part) << 31);
            }
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                const uint32_t b0 = home[i] & (S - 2);
                const uint32_t b2 = (b0 + 2) & (S - 1), b3 = (b0 + 3) & (S - 1);
                const ulonglong2 w0 = *reinterpret_cast<const ulonglong2*>(okey + b0);
                const ulonglong2 w1 = *reinterpret_cast<const ulonglong2*>(okey + b2);
                int h = w1.y == key[i] ? (int)b3 : -1;
                h = w1.x == key[i] ? (int)b2 : h;
                h = w0.y == key[i] ? (int)(b0 + 1) : h;
                h = w0.x == key[i] ? (int)b0 : h;
                hit[i] = h;
                // no match: a claim starts at the window's first empty slot from home on (slot
                // b0 lies before an odd home), or just past the window when it has none
                uint32_t f = (b0 + 4) & (S - 1);
                bool fr = false;
                if (w1.y == kEmpty) { f = b3; fr = true; }
                if (w1.x == kEmpty) { f = b2; fr = true; }
                if (w0.y == kEmpty) { f = b0 + 1; fr = true; }
                if (!(home[i] & 1) && w0.x == kEmpty) { f = b0; fr = true; }
                from[i] = f;
                free_at |= (uint32_t)fr << i;
            }
#pragma unroll
            for (int i = 0; i < kP2Rows; i++) {
                if ((j0 + i * 64 + lane >= cnt) | (home[i] >> 31)) continue;
                Part pr;
                if constexpr (HAS_V) pr = row_part<VT>(val[i], false); else pr = row_part<NoValue>(NoValue{}, false);
                int s = hit[i];
                if (s < 0) s = claim_from(key[i], from[i], (free_at >> i) & 1);
                if (s >= 0) add(s, 1, 1, pr.sum, pr.mn, pr.mx);
                else g_set_lost(g);
            }
        }
    }
    __syncthreads();
    // every key goes back (a claimed key may carry only NULL-value rows); states where a slot
    // took rows
    // (kFill slots per thread at a time: every global state they need is loaded before any
    // is stored)
    for (uint32_t s0 = tid; s0 < S; s0 += kFill * bd) {
        uint32_t c[kFill], r[kFill];
        unsigned long long gall[kFill], gcnt[kFill], gsum[kFill], gmn[kFill], gmx[kFill];
#pragma unroll
        for (int u = 0; u < kFill; u++) {
            const uint32_t s = s0 + u * bd;
            const bool in = s < S;
            c[u] = in ? ocnt[s] : 0u;
            r[u] = in && oall ? oall[s] : 0u;
            const int64_t gs = base + (in ? s : 0u);
            gall[u] = r[u] ? g.cntall[gs] : 0ull;
            gcnt[u] = c[u] ? g.cnt[gs] : 0ull;
            gsum[u] = c[u] && osum ? g.sum[gs] : 0ull;
            gmn[u] = c[u] && omin ? g.mn[gs] : 0ull;
            gmx[u] = c[u] && omax ? g.mx[gs] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kFill; u++) {
            const uint32_t s = s0 + u * bd;
            if (s >= S) continue;
            const int64_t gs = base + s;
            g.keys[gs] = okey[s];
            if (r[u]) g.cntall[gs] = gall[u] + r[u];
            if (!c[u]) continue;
            g.cnt[gs] = gcnt[u] + c[u];
            if (osum) {
                if (REAL)
                    g.sum[gs] = (unsigned long long)__double_as_longlong(
                        __longlong_as_double((long long)gsum[u]) + __longlong_as_double((long long)osum[s]));
                else
                    g.sum[gs] = gsum[u] + osum[s];
            }
            if (omin && omin[s] < gmn[u]) g.mn[gs] = omin[s];
            if (omax && omax[s] > gmx[u]) g.mx[gs] = omax[s];
        }
    }
    // NULL key and key == kEmpty: side slots, summed over the phase-1 workgroups (a strided
    // share per thread, then a sum per wavefront added with device atomics)
    if (part == 0) {
        for (int side = 0; side < 2; side++) {
            uint64_t rows = 0, cnt = 0, isum = 0, mn = ~0ull, mx = 0;
            double dsum = 0.0;
            for (int w = tid; w < a.W; w += bd) {
                const int64_t at = (int64_t)w * stride_k + a.kmax + side;
                rows += a.p_rows[at];
                cnt += a.p_cnt[at];
                if (osum) {
                    if (REAL) dsum += __longlong_as_double((long long)a.p_sum[at]);
                    else isum += a.p_sum[at];
                }
                if (omin) mn = min(mn, (uint64_t)a.p_min[at]);
                if (omax) mx = max(mx, (uint64_t)a.p_max[at]);
            }
            for (int d = 32; d >= 1; d >>= 1) {
                rows += __shfl_xor(rows, d, 64);
                cnt += __shfl_xor(cnt, d, 64);
                isum += __shfl_xor(isum, d, 64);
                dsum += __shfl_xor(dsum, d, 64);
                mn = min(mn, (uint64_t)__shfl_xor(mn, d, 64));
                mx = max(mx, (uint64_t)__shfl_xor(mx, d, 64));
            }
            if ((tid & 63) == 0 && rows) {
                const uint64_t gs = g.cap + side;
                if (g_claim_side(g, gs) && side == 1) nnew++;
                Part q;
                q.cntall = rows;
                q.cnt = cnt;
                q.sum = REAL ? (uint64_t)__double_as_longlong(dsum) : isum;
                q.mn = mn;
                q.mx = mx;
                g_accum_atomic(g, (int64_t)gs, q, fields);
            }
        }
    }
    // new keys: a sum per wavefront, counted once per wavefront
    for (int d = 32; d >= 1; d >>= 1) nnew += __shfl_xor(nnew, d, 64);
    if ((tid & 63) == 0) g_count_new(g, nnew);
}

// ---- launch helpers ----------------------------------------------------------------------

// Shift giving a key hash's phase-1 partition among P (a power of two): h >> shift, 64 for P = 1.
int32_t rec_shift_of(int32_t P) {
    int32_t bits = 0;
    while ((1 << bits) < P) bits++;
    return 64 - bits;
}

int grid_cap(int64_t items, int per, int cap) {
    int64_t g = (items + per - 1) / per;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

template <typename VT, int MODE>
int launch_sample(TwoPhaseBufs& b, const KeySpec& ks, const KeyCols& kc, const void* vals,
                  const uint64_t* vnull, const int8_t* mask, const PredSet& ps, uint32_t vpred,
                  int64_t n, int64_t nsamples, int64_t stride, uint64_t smask, hipStream_t s) {
    k_tp_sample<VT, MODE><<<grid_cap(nsamples, 1024, device_info().cus), kSampleBlock, 0, s>>>(
        ks, kc, (const VT*)vals, vnull, mask, ps, vpred, n, stride, nsamples, b.sample_keys,
        b.sample_cnt, smask);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

template <typename VT>
int launch_sample_m(int mode, TwoPhaseBufs& b, const KeySpec& ks, const KeyCols& kc,
                    const void* vals, const uint64_t* vnull, const int8_t* mask, const PredSet& ps,
                    uint32_t vpred, int64_t n, int64_t ns, int64_t st, uint64_t sm, hipStream_t s) {
    switch (mode) {
    case 0: return launch_sample<VT, 0>(b, ks, kc, vals, vnull, mask, ps, vpred, n, ns, st, sm, s);
    case 1: return launch_sample<VT, 1>(b, ks, kc, vals, vnull, mask, ps, vpred, n, ns, st, sm, s);
    case 2: return launch_sample<VT, 2>(b, ks, kc, vals, vnull, mask, ps, vpred, n, ns, st, sm, s);
    case 4: return launch_sample<VT, 4>(b, ks, kc, vals, vnull, mask, ps, vpred, n, ns, st, sm, s);
    default:
        if constexpr (VTraits<VT>::has)
            return launch_sample<VT, 3>(b, ks, kc, vals, vnull, mask, ps, vpred, n, ns, st, sm, s);
        return fail(QK_ERROR, "value range filter without a value column");
    }
}

P1Args p1_args(const TwoPhaseBufs& b) {
    P1Args a;
    a.image = b.p1_image;
    a.slot = b.p1_slot;
    a.meta = b.heavy_meta;
    a.kmax = b.kmax;
    a.H = b.H1;
    a.P = b.P;
    a.cap = b.cap;
    a.p_cnt = b.p_cnt;
    a.p_rows = b.p_rows;
    a.p_sum = b.p_sum;
    a.p_min = b.p_min;
    a.p_max = b.p_max;
    a.r_rec = b.r_rec;
    a.r_count = b.r_count;
    a.zero = b.zero_word;
    a.part_shift = rec_shift_of(b.P);
    a.accumulate = b.pending_rows > 0 ? 1 : 0;
    return a;
}

// gfx950 launches up to 160 KB of LDS per workgroup without a function attribute
template <typename VT, int MODE>
int launch_p1(TwoPhaseBufs& b, const KeySpec& ks, const KeyCols& kc, const void* vals, const uint64_t* vnull,
              const int8_t* mask, const PredSet& ps, uint32_t vpred, int64_t n, uint32_t fields, hipStream_t s) {
    const size_t lds = p1_lds_bytes(b.H1, b.P, b.vbytes, fields);
    k_tp_phase1<VT, MODE, void, kFieldsAny><<<b.W, b.block, lds, s>>>(b.gview, ks, kc, (const VT*)vals, vnull,
                                                                      mask, ps, vpred, n, fields, p1_args(b));
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

template <typename VT, int MODE, typename KT, int FP>
int launch_p1s(TwoPhaseBufs& b, const KeyCols& kc, const void* vals, const uint64_t* vnull, const PredSet& ps,
               int64_t n, uint32_t fields, hipStream_t s) {
    const size_t lds = p1_lds_bytes(b.H1, b.P, b.vbytes, fields);
    k_tp_phase1s<VT, MODE, KT, FP><<<b.W, b.block, lds, s>>>(b.gview, kc, (const VT*)vals, vnull, ps, (int32_t)n,
                                                             fields, p1_args(b));
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

template <typename VT, typename KT, int FP>
int launch_p1s_m(int mode, TwoPhaseBufs& b, const KeyCols& kc, const void* vals, const uint64_t* vnull,
                 const PredSet& ps, int64_t n, uint32_t fields, hipStream_t s) {
    if (mode == 0) return launch_p1s<VT, 0, KT, FP>(b, kc, vals, vnull, ps, n, fields, s);
    if (mode == 4) return launch_p1s<VT, 4, KT, FP>(b, kc, vals, vnull, ps, n, fields, s);
    if constexpr (VTraits<VT>::has) return launch_p1s<VT, 3, KT, FP>(b, kc, vals, vnull, ps, n, fields, s);
    return fail(QK_ERROR, "value range filter without a value column");
}

// Blocks of at most this many rows take the streaming phase 1 (its loads use 32-bit byte
// offsets: 8-byte columns and the prefetch past the end stay below 2^31 bytes).
constexpr int64_t kStreamMaxRows = (int64_t)1 << 27;

// The streaming form for one int64 or int32 key column, or two non-NULL int32 key columns
// packed into one word, under no WHERE or a folded one; the general form reads any key layout
// through the key spec and evaluates any WHERE.
template <typename VT>
int launch_p1_k(int mode, TwoPhaseBufs& b, const KeySpec& ks, const KeyCols& kc, const void* vals,
                const uint64_t* vnull, const int8_t* mask, const PredSet& ps, uint32_t vpred, int64_t n,
                uint32_t fields, hipStream_t s) {
    const bool sum_only = (fields & F_SUM) && !(fields & (F_MIN | F_MAX));
    const bool streamed = (mode == 0 || mode == 3 || mode == 4) && n <= kStreamMaxRows;
    b.last_form = 1;
    if (streamed && ks.single && ks.nkeys == 1 && ks.type[0] == QK_LONG) {
        if (sum_only) return launch_p1s_m<VT, int64_t, kFieldsSum>(mode, b, kc, vals, vnull, ps, n, fields, s);
        return launch_p1s_m<VT, int64_t, kFieldsAny>(mode, b, kc, vals, vnull, ps, n, fields, s);
    }
    if (streamed && ks.single && ks.nkeys == 1 && ks.type[0] == QK_INT) {
        if (sum_only) return launch_p1s_m<VT, int32_t, kFieldsSum>(mode, b, kc, vals, vnull, ps, n, fields, s);
        return launch_p1s_m<VT, int32_t, kFieldsAny>(mode, b, kc, vals, vnull, ps, n, fields, s);
    }
    if (streamed && !ks.single && ks.nkeys == 2 && ks.kw == 1 && ks.type[0] == QK_INT && ks.type[1] == QK_INT &&
        ks.word[0] == 0 && ks.shift[0] == 0 && ks.word[1] == 0 && ks.shift[1] == 32 && !kc.nm[0] && !kc.nm[1])
        return launch_p1s_m<VT, KeyPairI32, kFieldsAny>(mode, b, kc, vals, vnull, ps, n, fields, s);
    b.last_form = 0;
    switch (mode) {
    case 0: return launch_p1<VT, 0>(b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
    case 1: return launch_p1<VT, 1>(b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
    case 2: return launch_p1<VT, 2>(b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
    case 4: return launch_p1<VT, 4>(b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
    default:
        if constexpr (VTraits<VT>::has) return launch_p1<VT, 3>(b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
        return fail(QK_ERROR, "value range filter without a value column");
    }
}

template <typename VT>
int launch_p2(TwoPhaseBufs& b, const GView& g, uint32_t fields, hipStream_t s) {
    P2Args a;
    a.heavy = b.heavy_keys;
    a.meta = b.heavy_meta;
    a.kmax = b.kmax;
    a.H2 = b.H2;
    a.P = b.P;
    a.cap = b.cap;
    a.W = b.W;
    a.p_cnt = b.p_cnt;
    a.p_rows = b.p_rows;
    a.p_sum = b.p_sum;
    a.p_min = b.p_min;
    a.p_max = b.p_max;
    a.r_rec = b.r_rec;
    a.r_count = b.r_count;
    const bool sum_only = (fields & F_SUM) && !(fields & (F_MIN | F_MAX));
    const size_t olds = p2o_lds_bytes((int64_t)g.sub_mask + 1, fields);
    // sub-tables per partition: a power of two (both counts are), dealt to the dies by eights
    const int64_t F = (int64_t)(g.cap >> g.sub_bits);
    a.split = 0;
    while (((int64_t)b.P << a.split) < F) a.split++;
    // (the owned kernel addresses a run of regions of one partition with 32-bit offsets)
    const bool split_ok = (a.split == 0 || (a.split <= kMaxSplitBits && b.P % 8 == 0)) &&
                          (int64_t)b.P * b.cap * 8 < ((int64_t)1 << 31);
    if (((int64_t)b.P << a.split) == F && split_ok &&
        olds <= std::min<size_t>(device_info().lds_per_block, 160 * 1024)) {
        // the sub-tables fit the LDS: each workgroup folds into its own sub-table directly
        if (sum_only) k_tp_phase2_owned<VT, kFieldsSum><<<(int)F, b.block, olds, s>>>(g, fields, a);
        else k_tp_phase2_owned<VT, kFieldsAny><<<(int)F, b.block, olds, s>>>(g, fields, a);
    } else if (sum_only) {
        k_tp_phase2<VT, kFieldsSum><<<b.P, b.block, p2_lds_bytes(b.H2, b.W, fields), s>>>(g, fields, a);
    } else {
        k_tp_phase2<VT, kFieldsAny><<<b.P, b.block, p2_lds_bytes(b.H2, b.W, fields), s>>>(g, fields, a);
    }
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

}  // namespace

bool twophase_supported(int32_t kw, int32_t value_type) {
    (void)value_type;
    return kw == 1;
}

void twophase_free(TwoPhaseBufs& b) {
    void* ptrs[] = {b.heavy_keys, b.heavy_cnt, b.cand_cnt, b.heavy_meta, b.sample_keys, b.sample_cnt, b.hist, b.cand,
                    b.p1_image, b.p1_slot,
                    b.p_cnt, b.p_rows, b.p_sum, b.p_min, b.p_max, b.r_rec, b.r_count,
                    b.zero_word, b.gview};
    for (void* p : ptrs)
        if (p) pool_free(p);
    TwoPhaseBufs fresh;
    fresh.defer_rows = b.defer_rows;  // a setting, not a buffer
    b = fresh;
}

namespace {

// Rows the regions are sized for: one block, or several when phase 2 is deferred.
int64_t region_rows_for(const TwoPhaseBufs& b, int64_t n, bool defer) {
    if (!defer || b.defer_rows == 0) return n;
    const int64_t want = b.defer_rows > 0 ? b.defer_rows : std::min<int64_t>(4 * n, (int64_t)1 << 28);
    return std::max<int64_t>(n, std::min<int64_t>(want, (int64_t)1 << 30));
}

}  // namespace

bool twophase_needs_realloc(const TwoPhaseBufs& b, int32_t parts, int32_t value_type,
                            bool has_values, int64_t n, bool defer) {
    const int vbytes = has_values ? type_size(value_type) : 0;
    return !b.r_rec || b.block_alloc < n || b.vbytes != vbytes || b.P != parts ||
           b.rows_alloc < region_rows_for(b, n, defer);
}

int twophase_prepare(TwoPhaseBufs& b, int32_t parts, uint32_t fields, int32_t value_type,
                     bool has_values, int64_t n, bool defer, hipStream_t s) {
    if (!twophase_needs_realloc(b, parts, value_type, has_values, n, defer)) return QK_SUCCESS;
    const DeviceInfo& d = device_info();
    const int vbytes = has_values ? type_size(value_type) : 0;
    const int64_t rows = region_rows_for(b, n, defer);
    // the old buffers go back to the shared pool, where another stream or thread may be handed
    // them at once: wait for every kernel that may still read them (a deferred phase 2, on
    // whatever stream the table was last used from) before giving them up
    if (b.r_rec) QK_HIP(hipDeviceSynchronize());
    twophase_free(b);
    b.W = d.cus > 0 ? std::min(d.cus, 1024) : 256;
    b.P = parts;
    b.vbytes = vbytes;
    b.block = kBlock;
    // static LDS of the kernels (a few words) stays outside this budget
    const size_t lds = std::min<size_t>(d.lds_per_block, 160 * 1024) - 512;
    // phase-1 LDS: buckets of kBucket slots, at most half of the slots hold heavy keys
    int32_t H1 = 0;
    for (int32_t h = kBucket; ; h += kBucket) {
        if (p1_lds_bytes(h, parts, vbytes, fields) > lds) break;
        H1 = h;
    }
    b.H1 = H1;
    b.kmax = std::min<int32_t>(H1 / 2, 4096);
    int32_t H2 = 64;
    while (p2_lds_bytes(H2 + 64, b.W, fields) <= lds) H2 += 64;
    b.H2 = H2;
    // the heavy keys of one partition must fit the phase-2 table below its 4/5 fill limit
    b.kmax = std::min<int32_t>(b.kmax, b.H2 * 3 / 5);
    // regions: the mean rows per (workgroup, partition) and half again, plus a margin; rows
    // beyond a full region take the atomic path, so this bounds speed, not correctness
    const int64_t mean = (rows + (int64_t)b.W * parts - 1) / ((int64_t)b.W * parts);
    b.cap = (int32_t)((mean + mean / 2 + 96 + 31) / 32 * 32);
    b.rows_alloc = rows;
    b.block_alloc = n;
    const int64_t sslots = (int64_t)1 << (b.sample_bits + 1);
    const int64_t kp = b.kmax + 2;
    const int64_t regions = (int64_t)b.W * parts * b.cap;
    int st = QK_SUCCESS;
    auto A = [&](void** ptr, size_t bytes) {
        if (st == QK_SUCCESS) st = pool_alloc(ptr, bytes);
    };
    A((void**)&b.heavy_keys, (size_t)b.kmax * 8);
    A((void**)&b.heavy_cnt, (size_t)b.kmax * 4);
    A((void**)&b.cand_cnt, (size_t)b.kmax * 4);
    A((void**)&b.heavy_meta, (size_t)(parts + 8) * 4);
    A((void**)&b.sample_keys, (size_t)sslots * 8);
    A((void**)&b.sample_cnt, (size_t)sslots * 4);
    A((void**)&b.hist, (size_t)kHistBins * 4);
    A((void**)&b.cand, (size_t)b.kmax * 8);
    A((void**)&b.p1_image, (size_t)b.H1 * 8);
    A((void**)&b.p1_slot, (size_t)b.kmax * 4);
    A((void**)&b.p_cnt, (size_t)b.W * kp * 4);
    A((void**)&b.p_rows, (size_t)b.W * kp * 4);
    if (fields & F_SUM) A((void**)&b.p_sum, (size_t)b.W * kp * 8);
    if (fields & F_MIN) A((void**)&b.p_min, (size_t)b.W * kp * 8);
    if (fields & F_MAX) A((void**)&b.p_max, (size_t)b.W * kp * 8);
    A((void**)&b.r_rec, (size_t)regions * (8 + (size_t)vbytes));
    A((void**)&b.r_count, (size_t)b.W * parts * 4);
    A((void**)&b.zero_word, 64);
    A((void**)&b.gview, sizeof(GView));
    if (st != QK_SUCCESS) {
        twophase_free(b);
        return st;
    }
    // no heavy set until the next sample: K = 0, every partition empty
    QK_HIP(hipMemsetAsync(b.heavy_meta, 0, (size_t)(parts + 8) * 4, s));
    QK_HIP(hipMemsetAsync(b.zero_word, 0, 64, s));
    // the image of the empty heavy set
    k_tp_order<<<1, 1024, (size_t)b.H1 * 8 + (size_t)b.kmax * 12, s>>>(
        b.cand, b.cand_cnt, b.kmax, b.P, rec_shift_of(b.P), b.heavy_meta, b.heavy_keys, b.heavy_cnt, b.H1,
        b.p1_image, b.p1_slot);
    QK_HIP(hipGetLastError());
    b.heavy_ready = false;
    return QK_SUCCESS;
}

int twophase_sample(TwoPhaseBufs& b, const GView& g, const KeySpec& ks, const KeyCols& kc,
                    int32_t vtype, const void* vals, const uint64_t* vnull, const int8_t* mask,
                    const PredSet& ps, uint32_t vpred, int64_t n, int mode, hipStream_t s) {
    const int64_t sslots = (int64_t)1 << (b.sample_bits + 1);
    const int64_t nsamples = std::min<int64_t>(n, (int64_t)1 << b.sample_bits);
    const int64_t stride = std::max<int64_t>(1, n / nsamples);
    const int cus = device_info().cus;
    k_tp_clear_sample<<<grid_cap(sslots, 256, cus * 8), 256, 0, s>>>(b.sample_keys, b.sample_cnt,
                                                                     (uint64_t)sslots, b.hist, b.heavy_meta);
    QK_HIP(hipGetLastError());
    int st;
    if (!vals) {
        st = launch_sample_m<NoValue>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, nsamples, stride, sslots - 1, s);
    } else {
        switch (vtype) {
        case QK_INT: st = launch_sample_m<int32_t>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, nsamples, stride, sslots - 1, s); break;
        case QK_LONG: st = launch_sample_m<int64_t>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, nsamples, stride, sslots - 1, s); break;
        case QK_FLOAT: st = launch_sample_m<float>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, nsamples, stride, sslots - 1, s); break;
        default: st = launch_sample_m<double>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, nsamples, stride, sslots - 1, s); break;
        }
    }
    if (st != QK_SUCCESS) return st;
    k_tp_hist<<<kHistGrid, kHistBlock, 0, s>>>(b.sample_keys, b.sample_cnt, (uint64_t)sslots, b.hist);
    k_tp_threshold<<<1, kHistBins, 0, s>>>(b.hist, b.kmax, b.heavy_meta);
    k_tp_collect<<<kCollectGrid, kCollectBlock, 0, s>>>(b.sample_keys, b.sample_cnt, (uint64_t)sslots,
                                                               b.kmax, b.heavy_meta, b.cand, b.cand_cnt);
    k_tp_order<<<1, 1024, (size_t)b.H1 * 8 + (size_t)b.kmax * 12, s>>>(
        b.cand, b.cand_cnt, b.kmax, b.P, rec_shift_of(b.P), b.heavy_meta, b.heavy_keys, b.heavy_cnt, b.H1,
        b.p1_image, b.p1_slot);
    QK_HIP(hipGetLastError());
    b.heavy_ready = true;
    return QK_SUCCESS;
}


int twophase_phase1(TwoPhaseBufs& b, const GView& g, const KeySpec& ks, const KeyCols& kc,
                    int32_t vtype, const void* vals, const uint64_t* vnull, const int8_t* mask,
                    const PredSet& ps, uint32_t vpred, int64_t n, uint32_t fields, int mode,
                    hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (b.pending_rows > 0 && (vals ? vtype : 0) != b.pend_vtype)
        return fail(QK_ERROR, "group by: value column type changed between blocks");
    if (b.pending_rows + n > b.rows_alloc || n > b.block_alloc)
        return fail(QK_ERROR, "group by: two-phase buffers too small for the block");
    if (g.part_shift < 33 || g.part_shift > 64 || b.P > kMaxRecParts || (b.P & (b.P - 1)) ||
        (int64_t)(g.cap >> g.sub_bits) < b.P)
        return fail(QK_ERROR, "group by: unsupported partition count for the two-phase path");
    if (!b.gview_valid || memcmp(&b.gview_host, &g, sizeof(GView)) != 0) {
        b.gview_host = g;
        QK_HIP(hipMemcpyAsync(b.gview, &b.gview_host, sizeof(GView), hipMemcpyHostToDevice, s));
        b.gview_valid = true;
    }
    if (ev0) QK_HIP(hipEventRecord(ev0, s));
    int st;
    if (!vals) {
        st = launch_p1_k<NoValue>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s);
    } else {
        switch (vtype) {
        case QK_INT: st = launch_p1_k<int32_t>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s); break;
        case QK_LONG: st = launch_p1_k<int64_t>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s); break;
        case QK_FLOAT: st = launch_p1_k<float>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s); break;
        default: st = launch_p1_k<double>(mode, b, ks, kc, vals, vnull, mask, ps, vpred, n, fields, s); break;
        }
    }
    if (st != QK_SUCCESS) return st;
    if (ev1) QK_HIP(hipEventRecord(ev1, s));
    b.pending_rows += n;
    b.pend_vtype = vals ? vtype : 0;
    return QK_SUCCESS;
}

int twophase_flush(TwoPhaseBufs& b, const GView& g, uint32_t fields, hipStream_t s) {
    if (b.pending_rows == 0) return QK_SUCCESS;
    int st;
    switch (b.pend_vtype) {
    case 0: st = launch_p2<NoValue>(b, g, fields, s); break;
    case QK_INT: st = launch_p2<int32_t>(b, g, fields, s); break;
    case QK_LONG: st = launch_p2<int64_t>(b, g, fields, s); break;
    case QK_FLOAT: st = launch_p2<float>(b, g, fields, s); break;
    default: st = launch_p2<double>(b, g, fields, s); break;
    }
    if (st != QK_SUCCESS) return st;
    b.pending_rows = 0;
    return QK_SUCCESS;
}

}  // namespace qk
