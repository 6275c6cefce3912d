// batch.hip - many independent filtered GROUP BY queries over one device-resident table in
// one persistent kernel (qk_batch_* in include/qikkdb_amd.h).
//
// Query i reads rows [row0_i, row0_i + rows_i) of columns a, b (int32), ts and value (int64),
// keeps rows with lo_i <= ts < hi_i and groups them by key = (a * b_range + b) mod key_mod_i,
// a dense range [0, key_mod_i).  Its rows are cut into work units of `tile_rows` rows; the
// units of all queries form one range that persistent workgroups claim from a device counter.
//
// A unit of a query with at most `lds_groups` groups aggregates in a dense LDS table (count,
// sum, min, max per group) and then adds every touched group to the query's result with
// device atomics.  A unit of a larger query updates the result with device atomics per row
// (MIN/MAX only when they improve).  Results are dense arrays indexed by out_i + key; MIN/MAX
// are kept in an order-preserving unsigned encoding during the run and decoded at its end.
//
// The run is init (clear results and counters) -> persistent kernel -> finalize (decode),
// all reading the batch's size from a device header so that one captured hipGraph serves
// every batch that fits the engine's buffers.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace qk {
namespace {

constexpr int kBThreads = 1024;
constexpr int kBRows = 4;  // rows per thread per step of a unit

struct BQuery {
    int64_t row0, rows;   // table rows of the query
    int64_t unit0;        // first work unit
    int64_t ts_lo, ts_hi;
    int64_t out;          // offset of group 0 in the result arrays
    uint64_t magic;       // ~0 / key_mod + 1 (remainder by multiplication)
    int32_t mod;
    int32_t units;
};

struct BHeader {
    int32_t nq;
    int32_t pad;
    int64_t units;
    int64_t out_total;
};

struct BArgs {
    const int32_t* a;
    const int32_t* b;
    const int64_t* ts;
    const int64_t* v;
    int32_t b_range;
    int32_t tile_rows;
    int32_t lds_groups;
    const BHeader* hdr;
    const BQuery* q;
    unsigned long long* next_unit;
    uint32_t* done;             // [nq] finished units
    unsigned long long* comp;   // [nq] completion clock
    unsigned long long* t0;     // start clock
    unsigned long long* cnt;
    unsigned long long* sum;
    unsigned long long* mn;
    unsigned long long* mx;
};

// x mod d for 0 <= x < 2^32 and 1 <= d < 2^32 (Lemire, Kaser and Kurz, "Faster remainder by
// direct computation", 2019); other x take the division.
__device__ __forceinline__ uint32_t key_of(int32_t a, int32_t b, int32_t b_range, const BQuery& q) {
    const int64_t x = (int64_t)a * b_range + b;
    if ((uint64_t)x < (1ull << 32)) return (uint32_t)__umul64hi(q.magic * (uint64_t)x, (uint64_t)q.mod);
    int64_t r = x % q.mod;
    if (r < 0) r += q.mod;
    return (uint32_t)r;
}

__global__ void __launch_bounds__(256) k_batch_init(BArgs p) {
    const BHeader h = *p.hdr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < (uint64_t)h.out_total; i += stride) {
        p.cnt[i] = 0;
        p.sum[i] = 0;
        p.mn[i] = ~0ull;
        p.mx[i] = 0;
    }
    for (uint64_t i = i0; i < (uint64_t)h.nq; i += stride) {
        p.done[i] = 0;
        p.comp[i] = 0;
    }
    if (i0 == 0) {
        *p.next_unit = 0;
        *p.t0 = wall_clock64();
    }
}

__global__ void __launch_bounds__(256) k_batch_finalize(BArgs p) {
    const BHeader h = *p.hdr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)h.out_total; i += stride) {
        p.mn[i] = (unsigned long long)dec_i64(p.mn[i]);
        p.mx[i] = (unsigned long long)dec_i64(p.mx[i]);
    }
}

// A value every lane holds the same copy of, moved to scalar registers so that branches on it
// are uniform: a loop whose exit depends on an LDS value is otherwise compiled as a divergent
// loop, and lanes of one wave could then pass a barrier on different iterations.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ void add_global(const BArgs& p, int64_t at, uint32_t c, int64_t s, uint64_t mn,
                                           uint64_t mx) {
    atomicAdd(p.cnt + at, (unsigned long long)c);
    atomicAdd(p.sum + at, (unsigned long long)s);
    if (mn < *(volatile unsigned long long*)(p.mn + at)) atomicMin(p.mn + at, (unsigned long long)mn);
    if (mx > *(volatile unsigned long long*)(p.mx + at)) atomicMax(p.mx + at, (unsigned long long)mx);
}

__global__ void __launch_bounds__(kBThreads) k_batch_run(BArgs p) {
    extern __shared__ __align__(16) uint64_t lds[];
    __shared__ int64_t s_unit;
    __shared__ int32_t s_q;
    const int tid = threadIdx.x, bd = blockDim.x;
    const BHeader h = *p.hdr;
    const int G = p.lds_groups;
    int64_t* lsum = (int64_t*)lds;
    uint64_t* lmn = (uint64_t*)(lsum + G);
    uint64_t* lmx = lmn + G;
    uint32_t* lcnt = (uint32_t*)(lmx + G);
    int32_t q = 0;  // units are claimed in increasing order, so a workgroup's query index only grows
    int32_t prev_q = -1, prev_units = 0;
    while (true) {
        if (tid == 0) {
            // the previous unit's updates were issued before the barrier that ended it
            if (prev_q >= 0) {
                const uint32_t d = atomicAdd(p.done + prev_q, 1u) + 1;
                if (d == (uint32_t)prev_units) p.comp[prev_q] = wall_clock64();
            }
            const int64_t u = (int64_t)atomicAdd(p.next_unit, 1ull);
            s_unit = u;
            if (u < h.units) {
                while (q + 1 < h.nq && p.q[q + 1].unit0 <= u) q++;
                s_q = q;
            }
        }
        __syncthreads();
        const int64_t u = uniform64(s_unit);
        if (u >= h.units) break;  // every workgroup ends here once the counter passes the last unit
        q = (int32_t)uniform64(s_q);
        const BQuery Q = p.q[q];
        const int64_t r_begin = Q.row0 + (u - Q.unit0) * (int64_t)p.tile_rows;
        const int64_t r_end = min(r_begin + (int64_t)p.tile_rows, Q.row0 + Q.rows);
        const bool in_lds = Q.mod <= G;
        if (in_lds) {
            for (int g = tid; g < Q.mod; g += bd) {
                lsum[g] = 0;
                lmn[g] = ~0ull;
                lmx[g] = 0;
                lcnt[g] = 0;
            }
            __syncthreads();
        }
        for (int64_t base = r_begin; base < r_end; base += (int64_t)bd * kBRows) {
            int64_t tsv[kBRows];
#pragma unroll
            for (int k = 0; k < kBRows; k++) {
                const int64_t r = min(base + (int64_t)k * bd + tid, r_end - 1);
                tsv[k] = p.ts[r];
            }
#pragma unroll
            for (int k = 0; k < kBRows; k++) {
                const int64_t r = base + (int64_t)k * bd + tid;
                if (r >= r_end || tsv[k] < Q.ts_lo || tsv[k] >= Q.ts_hi) continue;
                const uint32_t key = key_of(p.a[r], p.b[r], p.b_range, Q);
                const int64_t val = p.v[r];
                if (in_lds) {
                    atomicAdd(&lcnt[key], 1u);
                    atomicAdd((unsigned long long*)&lsum[key], (unsigned long long)val);
                    const uint64_t e = enc_i64(val);
                    atomicMin((unsigned long long*)&lmn[key], (unsigned long long)e);
                    atomicMax((unsigned long long*)&lmx[key], (unsigned long long)e);
                } else {
                    const uint64_t e = enc_i64(val);
                    add_global(p, Q.out + key, 1u, val, e, e);
                }
            }
        }
        if (in_lds) {
            __syncthreads();
            for (int g = tid; g < Q.mod; g += bd)
                if (lcnt[g]) add_global(p, Q.out + g, lcnt[g], lsum[g], lmn[g], lmx[g]);
        }
        __syncthreads();  // the unit is done (and its LDS free) for every thread
        prev_q = q;
        prev_units = Q.units;
    }
    if (tid == 0 && prev_q >= 0) {
        const uint32_t d = atomicAdd(p.done + prev_q, 1u) + 1;
        if (d == (uint32_t)prev_units) p.comp[prev_q] = wall_clock64();
    }
}

__global__ void __launch_bounds__(256) k_key_column(const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                    int32_t b_range, BQuery q, int64_t n, int64_t* out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = key_of(a[i], b[i], b_range, q);
}

uint64_t magic_of(int32_t mod) { return ~0ull / (uint64_t)(uint32_t)mod + 1; }

}  // namespace
}  // namespace qk

using namespace qk;

struct qk_batch_s {
    qk_batch_columns cols{};
    qk_batch_tuning tune{};
    int device = 0;
    // device buffers (grown as needed)
    void* plan = nullptr;  // BHeader + BQuery[nq]
    size_t plan_bytes = 0;
    uint32_t* done = nullptr;
    unsigned long long* comp = nullptr;
    int32_t q_cap = 0;
    unsigned long long* counters = nullptr;  // [0] next unit, [1] start clock
    unsigned long long *cnt = nullptr, *sum = nullptr, *mn = nullptr, *mx = nullptr;
    int64_t out_cap = 0;
    void* host_plan = nullptr;  // pinned copy of the plan
    size_t host_plan_bytes = 0;
    std::vector<int64_t> offsets;
    int32_t last_nq = 0;
    // captured graph of init -> run -> finalize for the current buffers and tuning
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;
};

namespace qk {
namespace {

void free_graph(qk_batch_s* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_ok = false;
}

void free_buffers(qk_batch_s* b) {
    void* ptrs[] = {b->plan, b->done, b->comp, b->counters, b->cnt, b->sum, b->mn, b->mx};
    for (void* p : ptrs)
        if (p) pool_free(p);
    b->plan = nullptr;
    b->done = nullptr;
    b->comp = nullptr;
    b->counters = nullptr;
    b->cnt = b->sum = b->mn = b->mx = nullptr;
    b->plan_bytes = 0;
    b->q_cap = 0;
    b->out_cap = 0;
    if (b->host_plan) (void)hipHostFree(b->host_plan);
    b->host_plan = nullptr;
    b->host_plan_bytes = 0;
    free_graph(b);
}

int ensure_buffers(qk_batch_s* b, int32_t nq, int64_t out_total) {
    const size_t pbytes = sizeof(BHeader) + sizeof(BQuery) * (size_t)std::max(nq, 1);
    if (nq <= b->q_cap && out_total <= b->out_cap && pbytes <= b->plan_bytes && b->counters) return QK_SUCCESS;
    const int32_t qc = std::max(nq, std::max(b->q_cap * 2, 1024));
    const int64_t oc = std::max(out_total, std::max(b->out_cap + b->out_cap / 2, (int64_t)1 << 20));
    free_buffers(b);
    int st = QK_SUCCESS;
    auto A = [&](void** p, size_t bytes) {
        if (st == QK_SUCCESS) st = pool_alloc(p, bytes);
    };
    const size_t pb = sizeof(BHeader) + sizeof(BQuery) * (size_t)qc;
    A(&b->plan, pb);
    A((void**)&b->done, (size_t)qc * 4);
    A((void**)&b->comp, (size_t)qc * 8);
    A((void**)&b->counters, 64);
    A((void**)&b->cnt, (size_t)oc * 8);
    A((void**)&b->sum, (size_t)oc * 8);
    A((void**)&b->mn, (size_t)oc * 8);
    A((void**)&b->mx, (size_t)oc * 8);
    if (st != QK_SUCCESS) {
        free_buffers(b);
        return st;
    }
    if (hipHostMalloc(&b->host_plan, pb) != hipSuccess) {
        b->host_plan = nullptr;
        free_buffers(b);
        return fail(QK_OUT_OF_MEMORY, "qk_batch: pinned plan buffer");
    }
    b->host_plan_bytes = pb;
    b->plan_bytes = pb;
    b->q_cap = qc;
    b->out_cap = oc;
    return QK_SUCCESS;
}

BArgs args_of(const qk_batch_s* b) {
    BArgs p;
    p.a = b->cols.a;
    p.b = b->cols.b;
    p.ts = b->cols.ts;
    p.v = b->cols.value;
    p.b_range = b->cols.b_range;
    p.tile_rows = b->tune.tile_rows;
    p.lds_groups = b->tune.lds_groups;
    p.hdr = (const BHeader*)b->plan;
    p.q = (const BQuery*)((const char*)b->plan + sizeof(BHeader));
    p.next_unit = b->counters;
    p.t0 = b->counters + 1;
    p.done = b->done;
    p.comp = b->comp;
    p.cnt = b->cnt;
    p.sum = b->sum;
    p.mn = b->mn;
    p.mx = b->mx;
    return p;
}

size_t run_lds_bytes(const qk_batch_tuning& t) { return (size_t)t.lds_groups * 28 + 64; }

int launch_all(qk_batch_s* b, hipStream_t s) {
    const BArgs p = args_of(b);
    const int cus = device_info().cus;
    k_batch_init<<<cus * 8, 256, 0, s>>>(p);
    QK_HIP(hipGetLastError());
    k_batch_run<<<cus * b->tune.wgs_per_cu, kBThreads, run_lds_bytes(b->tune), s>>>(p);
    QK_HIP(hipGetLastError());
    k_batch_finalize<<<cus * 8, 256, 0, s>>>(p);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

// Capture init -> run -> finalize once for the current buffers and tuning.
int build_graph(qk_batch_s* b) {
    free_graph(b);
    hipStream_t cs = nullptr;
    QK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    int st = QK_SUCCESS;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) st = QK_RUNTIME_ERROR;
    if (st == QK_SUCCESS) {
        const int lst = launch_all(b, cs);
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (lst != QK_SUCCESS || e != hipSuccess) st = QK_RUNTIME_ERROR;
    }
    if (st == QK_SUCCESS && hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0) != hipSuccess) {
        b->graph = nullptr;
        st = QK_RUNTIME_ERROR;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    (void)hipGetLastError();
    b->graph_ok = st == QK_SUCCESS;
    return QK_SUCCESS;  // without a graph the kernels are launched one by one
}

int check_tuning(const qk_batch_tuning& t) {
    if (t.tile_rows < 4096 || t.tile_rows % 4096) return fail(QK_ERROR, "qk_batch: tile_rows must be a positive multiple of 4096");
    if (t.lds_groups < 0 || run_lds_bytes(t) > std::min<size_t>(device_info().lds_per_block, 160 * 1024) - 256)
        return fail(QK_ERROR, "qk_batch: lds_groups does not fit the LDS");
    if (t.wgs_per_cu < 1 || t.wgs_per_cu > 2) return fail(QK_ERROR, "qk_batch: wgs_per_cu must be 1 or 2");
    if (t.wgs_per_cu == 2 && run_lds_bytes(t) > std::min<size_t>(device_info().lds_per_cu, 160 * 1024) / 2 - 256)
        return fail(QK_ERROR, "qk_batch: two workgroups per CU need lds_groups below half the LDS");
    return QK_SUCCESS;
}

int run_batch(qk_batch_s* b, int32_t nq, const qk_batch_query* q, hipStream_t s) {
    const qk_batch_columns& c = b->cols;
    const int64_t nblocks = (c.rows + c.block_rows - 1) / c.block_rows;
    int64_t units = 0, out = 0;
    b->offsets.assign((size_t)nq + 1, 0);
    std::vector<BQuery> bq((size_t)nq);
    for (int32_t i = 0; i < nq; i++) {
        const qk_batch_query& x = q[i];
        if (x.key_mod < 1 || x.key_mod > (1 << 24)) return fail(QK_ERROR, "qk_batch: query %d: key_mod must be 1..2^24", i);
        if (x.first_block < 0 || x.nblocks < 1 || x.first_block + (int64_t)x.nblocks > nblocks)
            return fail(QK_ERROR, "qk_batch: query %d: block range outside the table", i);
        BQuery& d = bq[(size_t)i];
        d.row0 = (int64_t)x.first_block * c.block_rows;
        d.rows = std::min<int64_t>((int64_t)(x.first_block + x.nblocks) * c.block_rows, c.rows) - d.row0;
        d.unit0 = units;
        d.units = (int32_t)((d.rows + b->tune.tile_rows - 1) / b->tune.tile_rows);
        d.ts_lo = x.ts_lo;
        d.ts_hi = x.ts_hi;
        d.out = out;
        d.mod = x.key_mod;
        d.magic = magic_of(x.key_mod);
        units += d.units;
        b->offsets[(size_t)i] = out;
        out += x.key_mod;
    }
    b->offsets[(size_t)nq] = out;
    const bool grew = nq > b->q_cap || out > b->out_cap || !b->counters;
    QK_TRY(ensure_buffers(b, nq, out));
    // the plan travels through pinned memory: wait until the previous run has read it
    QK_HIP(hipStreamSynchronize(s));
    BHeader hd{};
    hd.nq = nq;
    hd.units = units;
    hd.out_total = out;
    memcpy(b->host_plan, &hd, sizeof hd);
    if (nq) memcpy((char*)b->host_plan + sizeof hd, bq.data(), sizeof(BQuery) * (size_t)nq);
    QK_HIP(hipMemcpyAsync(b->plan, b->host_plan, sizeof hd + sizeof(BQuery) * (size_t)nq, hipMemcpyHostToDevice, s));
    b->last_nq = nq;
    if (b->tune.use_graph) {
        if (grew || !b->graph_ok) QK_TRY(build_graph(b));
        if (b->graph_ok) {
            QK_HIP(hipGraphLaunch(b->graph, s));
            return QK_SUCCESS;
        }
    }
    return launch_all(b, s);
}

}  // namespace
}  // namespace qk

extern "C" {

int qk_batch_create(qk_batch* out, const qk_batch_columns* cols) {
    if (!out || !cols) return fail(QK_ERROR, "qk_batch_create: null argument");
    *out = nullptr;
    QK_TRY(ensure_init());
    if (!cols->a || !cols->b || !cols->ts || !cols->value) return fail(QK_ERROR, "qk_batch_create: null column");
    if (cols->rows < 0 || cols->block_rows < 1 || cols->b_range < 1)
        return fail(QK_ERROR, "qk_batch_create: rows >= 0, block_rows >= 1 and b_range >= 1 required");
    qk_batch_s* b = new (std::nothrow) qk_batch_s();
    if (!b) return fail(QK_OUT_OF_MEMORY, "qk_batch_create: host allocation");
    b->cols = *cols;
    b->tune.tile_rows = 1 << 18;
    b->tune.lds_groups = 4096;
    b->tune.wgs_per_cu = 1;
    b->tune.use_graph = 1;
    *out = b;
    return QK_SUCCESS;
}

int qk_batch_destroy(qk_batch b) {
    if (!b) return QK_SUCCESS;
    (void)hipDeviceSynchronize();
    free_buffers(b);
    delete b;
    return QK_SUCCESS;
}

int qk_batch_set_tuning(qk_batch b, const qk_batch_tuning* t) {
    if (!b || !t) return fail(QK_ERROR, "qk_batch_set_tuning: null argument");
    QK_TRY(check_tuning(*t));
    const bool same = memcmp(&b->tune, t, sizeof *t) == 0;
    b->tune = *t;
    if (!same) free_graph(b);
    return QK_SUCCESS;
}

int qk_batch_get_tuning(qk_batch b, qk_batch_tuning* t) {
    if (!b || !t) return fail(QK_ERROR, "qk_batch_get_tuning: null argument");
    *t = b->tune;
    return QK_SUCCESS;
}

int qk_batch_run(qk_batch b, int32_t nq, const qk_batch_query* q, void* stream) {
    if (!b || (nq > 0 && !q) || nq < 0) return fail(QK_ERROR, "qk_batch_run: bad argument");
    QK_TRY(ensure_init());
    return run_batch(b, nq, q, as_stream(stream));
}

int qk_batch_autotune(qk_batch b, int32_t nq, const qk_batch_query* q, qk_batch_tuning* out, void* stream) {
    if (!b || nq < 1 || !q) return fail(QK_ERROR, "qk_batch_autotune: bad argument");
    QK_TRY(ensure_init());
    hipStream_t s = as_stream(stream);
    const qk_batch_tuning keep = b->tune;
    std::vector<qk_batch_tuning> cands;
    for (int tile : {1 << 16, 1 << 18, 1 << 20})
        for (int groups : {2048, 4096, 5120})
            for (int wpc : {1, 2}) {
                qk_batch_tuning t{tile, groups, wpc, keep.use_graph};
                if (check_tuning(t) == QK_SUCCESS) cands.push_back(t);
            }
    (void)qk_last_error_message();
    hipEvent_t e0, e1;
    QK_HIP(hipEventCreate(&e0));
    QK_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    qk_batch_tuning best_t = keep;
    int st = QK_SUCCESS;
    for (const qk_batch_tuning& t : cands) {
        b->tune = t;
        free_graph(b);
        st = run_batch(b, nq, q, s);  // warm-up (and graph capture)
        if (st != QK_SUCCESS) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = QK_RUNTIME_ERROR; break; }
        st = run_batch(b, nq, q, s);
        if (st != QK_SUCCESS) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { st = QK_RUNTIME_ERROR; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            best_t = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tune = st == QK_SUCCESS ? best_t : keep;
    free_graph(b);
    if (st != QK_SUCCESS) return st;
    if (out) *out = best_t;
    return QK_SUCCESS;
}

int qk_batch_results(qk_batch b, const int64_t** count, const int64_t** sum, const int64_t** min,
                     const int64_t** max, const int64_t** offsets_host) {
    if (!b) return fail(QK_ERROR, "qk_batch_results: null batch");
    if (count) *count = (const int64_t*)b->cnt;
    if (sum) *sum = (const int64_t*)b->sum;
    if (min) *min = (const int64_t*)b->mn;
    if (max) *max = (const int64_t*)b->mx;
    if (offsets_host) *offsets_host = b->offsets.data();
    return QK_SUCCESS;
}

int qk_batch_completion_ms(qk_batch b, double* out, int32_t nq) {
    if (!b || !out || nq < 0 || nq > b->last_nq) return fail(QK_ERROR, "qk_batch_completion_ms: bad argument");
    QK_HIP(hipDeviceSynchronize());
    std::vector<unsigned long long> comp((size_t)nq);
    unsigned long long t0 = 0;
    if (nq) QK_HIP(hipMemcpy(comp.data(), b->comp, (size_t)nq * 8, hipMemcpyDeviceToHost));
    QK_HIP(hipMemcpy(&t0, b->counters + 1, 8, hipMemcpyDeviceToHost));
    int rate_khz = 0;
    QK_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
    const double per_ms = rate_khz > 0 ? (double)rate_khz : 100000.0;
    for (int32_t i = 0; i < nq; i++) out[i] = comp[(size_t)i] >= t0 ? (double)(comp[(size_t)i] - t0) / per_ms : 0.0;
    return QK_SUCCESS;
}

int qk_batch_key_column(const int32_t* a, const int32_t* b, int32_t b_range, int32_t key_mod, int32_t n,
                        int64_t* out, void* stream) {
    if (n < 0 || (n > 0 && (!a || !b || !out))) return fail(QK_ERROR, "qk_batch_key_column: bad argument");
    if (key_mod < 1 || key_mod > (1 << 24) || b_range < 1) return fail(QK_ERROR, "qk_batch_key_column: bad key_mod or b_range");
    QK_TRY(ensure_init());
    if (n == 0) return QK_SUCCESS;
    BQuery q{};
    q.mod = key_mod;
    q.magic = magic_of(key_mod);
    const int grid = (int)std::min<int64_t>(((int64_t)n + 255) / 256, (int64_t)device_info().cus * 16);
    k_key_column<<<grid, 256, 0, as_stream(stream)>>>(a, b, b_range, q, n, out);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

}  // extern "C"
