// batch.hip - many independent filtered GROUP BY queries over one device-resident table
// (qk_batch_* in include/qikkdb_amd.h).
//
// Query i reads rows [row0_i, row0_i + rows_i) of columns a, b (int32), ts and value (int64),
// keeps rows with lo_i <= ts < hi_i and groups them by key = (a * b_range + b) mod key_mod_i,
// a dense range [0, key_mod_i).  The table is cut into tiles of `tile_rows` rows; a work unit
// is one (tile, query) pair.  Units are listed tile by tile, so the workgroups running at one
// time read the same rows for different queries and most of those reads hit the L2 and the
// memory-side cache instead of HBM.  Persistent workgroups claim units from a device counter.
//
// Small queries (key_mod <= lds_groups) aggregate a unit in a dense LDS table and add the
// unit's groups to the result with device atomics.
//
// Large queries take two passes instead of per-row device atomics.  Pass 1 counts a unit's
// passing rows per key chunk (chunk = key >> chunk_bits of the query) in LDS, takes one slice
// of a scratch pool for all of them, and writes the rows (key, value) into the slice grouped
// by chunk, recording where each chunk starts.  Pass 2 gives each (query, chunk) to one
// workgroup, whose wavefronts each take one unit's segment of the chunk at a time (coalesced
// reads), aggregate it in a dense LDS table, and add the table to the result with plain
// read-modify-write (the workgroup is the only writer of those groups).  The chunk width of a
// query is chosen from its estimated rows so that no chunk holds much more than a fair share
// of a pass, and large queries are scheduled in waves whose estimated rows fit the pool; a
// unit whose rows do not fit takes the atomic path.  Rows per query are estimated by sampling
// before the run.
//
// Results are dense arrays indexed by out_i + key.  MIN/MAX are kept in an order-preserving
// unsigned encoding during the run and decoded at its end.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace qk {
namespace {

constexpr int kBThreads = 1024;
constexpr int kBRows = 4;            // rows per thread per step of a unit
constexpr int kSamples = 16384;      // ts samples per block for the row estimates
constexpr int kMaxKeyBits = 24;      // key_mod <= 2^24
constexpr int kMinChunkBits = 6;     // narrowest key chunk of a large query: 64 groups
constexpr int64_t kDefaultScratchRows = (int64_t)1 << 32;  // 48 GiB at most, allocated as needed
constexpr int kXcds = 8;             // dies of the device, each with its own L2
constexpr int kCtrStride = 16;       // counters per wave: kXcds unit lists, pass-2 units, pool top
constexpr int kCtrPass2 = kXcds;
constexpr int kCtrPool = kXcds + 1;

struct BQuery {
    int64_t row0, rows;   // table rows of the query
    int64_t tile0;        // tile holding row0
    int64_t ts_lo, ts_hi;
    int64_t out;          // offset of group 0 in the result arrays
    uint64_t magic;       // ~0 / key_mod + 1 (remainder by multiplication)
    int64_t desc;         // large: first chunk offset of its units (units x (chunks + 1) int32)
    int64_t segb;         // large: first slice base of its units (int64 per unit)
    int64_t chunk0;       // large: first pass-2 unit
    int32_t mod;
    int32_t units;        // tiles the query touches
    int32_t chunks;       // large: key chunks; 0 for a small query
    int32_t cbits;        // large: chunk = key >> cbits
    int32_t index;        // position in the caller's list
    int32_t pad;
};

struct BUnit {
    int32_t q;     // query (plan order)
    int32_t tile;  // table tile
};

struct BWave {
    int64_t unit_off[kXcds + 1];     // pass-1 units of list x: [unit_off[x], unit_off[x + 1])
    int64_t chunk_begin, chunk_end;  // pass-2 units
    int32_t q_begin, q_end;          // queries of the wave (plan order)
};

struct BHeader {
    int32_t nq, nwaves;
    int64_t out_total;
    int64_t scratch_cap;  // rows
    int64_t nunits;
};

struct BArgs {
    const int32_t* a;
    const int32_t* b;
    const int64_t* ts;
    const int64_t* v;
    int32_t b_range;
    int32_t tile_rows;
    int32_t lds_groups;
    int32_t chunk_max;          // widest key chunk of a large query (a power of two)
    const BHeader* hdr;
    const BWave* waves;
    const BQuery* q;
    const BUnit* units;
    unsigned long long* ctr;    // kCtrStride per wave
    unsigned long long* t0;     // start clock
    uint32_t* done;             // [nq] finished units (pass 1 for small, pass 2 for large queries)
    unsigned long long* comp;   // [nq, caller order] completion clock
    int32_t* desc;
    int64_t* segb;
    uint32_t* sk;               // scratch pool: keys
    int64_t* sv;                // scratch pool: values
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

__device__ __forceinline__ void mark_done(const BArgs& p, int32_t q, uint32_t of) {
    const uint32_t d = atomicAdd(p.done + q, 1u) + 1;
    if (d == of) p.comp[p.q[q].index] = wall_clock64();
}

// Exclusive prefix sum over a workgroup (blockDim a multiple of 64); *total gets the sum.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < nw; w++) {
        if (w < wid) before += wsum[w];
        all += wsum[w];
    }
    *total = all;
    __syncthreads();
    return before + x - v;
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
    for (uint64_t i = i0; i < (uint64_t)kCtrStride * h.nwaves; i += stride) p.ctr[i] = 0;
    if (i0 == 0) *p.t0 = wall_clock64();
}

__global__ void __launch_bounds__(256) k_batch_finalize(BArgs p) {
    const BHeader h = *p.hdr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)h.out_total; i += stride) {
        p.mn[i] = (unsigned long long)dec_i64(p.mn[i]);
        p.mx[i] = (unsigned long long)dec_i64(p.mx[i]);
    }
}

// ts samples of every block: out[blk * per + i] = ts[row of sample i of block blk], the
// samples spread evenly over the block.
__global__ void __launch_bounds__(256) k_batch_sample_ts(const int64_t* __restrict__ ts, int64_t rows,
                                                         int64_t block_rows, int32_t per, int64_t nblocks,
                                                         int64_t* __restrict__ out) {
    const int64_t n = nblocks * per;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t blk = i / per, j = i % per;
        const int64_t r0 = blk * block_rows, len = min(block_rows, rows - r0);
        out[i] = ts[r0 + j * len / per];
    }
}

// Claims the next pass-1 unit of a wave: first from the list of this workgroup's die (units
// of one tile share that die's L2), then from the other lists once it is empty.  -1 when
// every list is empty.  `xi` is the caller's position in that order.
__device__ __forceinline__ int64_t claim_unit(const BArgs& p, const BWave& W, int32_t wave, int& xi) {
    const int home = blockIdx.x % kXcds;
    while (xi < kXcds) {
        const int x = (home + xi) % kXcds;
        const int64_t n = W.unit_off[x + 1] - W.unit_off[x];
        if (n > 0) {
            const int64_t c = (int64_t)atomicAdd(p.ctr + kCtrStride * wave + x, 1ull);
            if (c < n) return W.unit_off[x] + c;
        }
        xi++;
    }
    return -1;
}

__global__ void __launch_bounds__(kBThreads) k_batch_pass1(BArgs p, int32_t wave) {
    extern __shared__ __align__(16) uint64_t lds[];
    __shared__ int64_t s_unit, s_base;
    __shared__ uint32_t wsum[16];
    const int tid = threadIdx.x, bd = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nwv = bd >> 6;
    const BHeader h = *p.hdr;
    const BWave W = p.waves[wave];
    const int G = p.lds_groups;
    int32_t prev_q = -1;
    uint32_t prev_units = 0;
    int xi = 0;
    while (true) {
        if (tid == 0) {
            if (prev_q >= 0) mark_done(p, prev_q, prev_units);
            s_unit = claim_unit(p, W, wave, xi);
        }
        __syncthreads();
        const int64_t u = uniform64(s_unit);
        if (u < 0) break;  // every workgroup ends here once all lists of the wave are empty
        const BUnit U = p.units[u];
        const int32_t q = (int32_t)uniform64(U.q);
        const int64_t tile = uniform64(U.tile);
        const BQuery Q = p.q[q];
        const int64_t k = tile - Q.tile0;  // the query's unit index
        const int64_t r_begin = max(Q.row0, tile * (int64_t)p.tile_rows);
        const int64_t r_end = min((tile + 1) * (int64_t)p.tile_rows, Q.row0 + Q.rows);
        prev_q = Q.chunks == 0 ? q : -1;  // a large query is done after its pass-2 chunks
        prev_units = (uint32_t)Q.units;
        if (Q.chunks == 0) {
            // small query: dense LDS table of key_mod groups
            int64_t* lsum = (int64_t*)lds;
            uint64_t* lmn = (uint64_t*)(lsum + G);
            uint64_t* lmx = lmn + G;
            uint32_t* lcnt = (uint32_t*)(lmx + G);
            for (int g = tid; g < Q.mod; g += bd) {
                lsum[g] = 0;
                lmn[g] = ~0ull;
                lmx[g] = 0;
                lcnt[g] = 0;
            }
            __syncthreads();
            // every column of a step is loaded at once (one wait per step); most of these
            // lines are read by other queries of the same tile anyway
            for (int64_t base = r_begin; base < r_end; base += (int64_t)bd * kBRows) {
                int64_t tsv[kBRows], vv[kBRows];
                int32_t av[kBRows], bv[kBRows];
#pragma unroll
                for (int k = 0; k < kBRows; k++) {
                    const int64_t r = min(base + (int64_t)k * bd + tid, r_end - 1);
                    tsv[k] = p.ts[r];
                    av[k] = p.a[r];
                    bv[k] = p.b[r];
                    vv[k] = p.v[r];
                }
#pragma unroll
                for (int k = 0; k < kBRows; k++) {
                    const int64_t r = base + (int64_t)k * bd + tid;
                    const bool pass = (r < r_end) & (tsv[k] >= Q.ts_lo) & (tsv[k] < Q.ts_hi);
                    if (!pass) continue;
                    const uint32_t key = key_of(av[k], bv[k], p.b_range, Q);
                    const uint64_t e = enc_i64(vv[k]);
                    atomicAdd(&lcnt[key], 1u);
                    atomicAdd((unsigned long long*)&lsum[key], (unsigned long long)vv[k]);
                    atomicMin((unsigned long long*)&lmn[key], (unsigned long long)e);
                    atomicMax((unsigned long long*)&lmx[key], (unsigned long long)e);
                }
            }
            __syncthreads();
            for (int g = tid; g < Q.mod; g += bd)
                if (lcnt[g]) add_global(p, Q.out + g, lcnt[g], lsum[g], lmn[g], lmx[g]);
        } else {
            // large query, pass 1: rows per chunk, one slice of the pool, rows written by chunk.
            // The first sweep keeps one pass bit per row in LDS (a wavefront ballot per 64
            // rows), so the second sweep does not read ts again.
            const int nch = Q.chunks;
            const int cbits = Q.cbits;
            uint64_t* mask = lds;  // tile_rows / 64 words
            uint32_t* hist = (uint32_t*)(mask + (p.tile_rows >> 6));
            uint32_t* cur = hist + nch;
            for (int c = tid; c < nch; c += bd) hist[c] = 0;
            __syncthreads();
            int word = wid;
            for (int64_t base = r_begin; base < r_end; base += (int64_t)bd * kBRows) {
                int64_t tsv[kBRows];
                int32_t av[kBRows], bv[kBRows];
#pragma unroll
                for (int j = 0; j < kBRows; j++) {
                    const int64_t r = min(base + (int64_t)j * bd + tid, r_end - 1);
                    tsv[j] = p.ts[r];
                    av[j] = p.a[r];
                    bv[j] = p.b[r];
                }
#pragma unroll
                for (int j = 0; j < kBRows; j++, word += nwv) {
                    const int64_t r = base + (int64_t)j * bd + tid;
                    const bool pass = (r < r_end) & (tsv[j] >= Q.ts_lo) & (tsv[j] < Q.ts_hi);
                    const uint64_t m = __ballot(pass);
                    if (lane == 0) mask[word] = m;
                    if (pass) atomicAdd(&hist[key_of(av[j], bv[j], p.b_range, Q) >> cbits], 1u);
                }
            }
            __syncthreads();
            // exclusive offsets: thread t owns chunks [t * per, (t + 1) * per)
            const int per = (nch + bd - 1) / bd;
            const int c0 = min(nch, tid * per), c1 = min(nch, c0 + per);
            uint32_t run = 0;
            for (int c = c0; c < c1; c++) run += hist[c];
            uint32_t total;
            uint32_t off = block_scan(run, wsum, &total);
            int32_t* d = p.desc + Q.desc + k * (int64_t)(nch + 1);
            for (int c = c0; c < c1; c++) {
                cur[c] = off;
                d[c] = (int32_t)off;
                off += hist[c];
            }
            if (tid == 0) {
                d[nch] = (int32_t)total;
                const int64_t base =
                    total ? (int64_t)atomicAdd(p.ctr + kCtrStride * wave + kCtrPool, (unsigned long long)total) : 0;
                s_base = (base + (int64_t)total <= h.scratch_cap) ? base : -1;
                p.segb[Q.segb + k] = s_base;
            }
            __syncthreads();
            const int64_t sbase = uniform64(s_base);
            word = wid;
            for (int64_t base = r_begin; base < r_end; base += (int64_t)bd * kBRows) {
                int32_t av[kBRows], bv[kBRows];
                int64_t vv[kBRows];
#pragma unroll
                for (int j = 0; j < kBRows; j++) {
                    const int64_t r = min(base + (int64_t)j * bd + tid, r_end - 1);
                    av[j] = p.a[r];
                    bv[j] = p.b[r];
                    vv[j] = p.v[r];
                }
#pragma unroll
                for (int j = 0; j < kBRows; j++, word += nwv) {
                    if (!((mask[word] >> lane) & 1)) continue;
                    const uint32_t key = key_of(av[j], bv[j], p.b_range, Q);
                    const int64_t val = vv[j];
                    if (sbase >= 0) {
                        const uint32_t at = atomicAdd(&cur[key >> cbits], 1u);
                        p.sk[sbase + at] = key;
                        p.sv[sbase + at] = val;
                    } else {
                        const uint64_t e = enc_i64(val);
                        add_global(p, Q.out + key, 1u, val, e, e);
                    }
                }
            }
        }
        __syncthreads();  // the unit is done (and its LDS free) for every thread
    }
    if (tid == 0 && prev_q >= 0) mark_done(p, prev_q, prev_units);
}

// Pass 2: one (large query, key chunk) per unit.
__global__ void __launch_bounds__(kBThreads) k_batch_pass2(BArgs p, int32_t wave) {
    extern __shared__ __align__(16) uint64_t lds[];
    __shared__ int64_t s_unit;
    __shared__ int32_t s_q;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int lane = tid & 63, wid = tid >> 6, nwv = bd >> 6;
    const BWave W = p.waves[wave];
    const int S = p.chunk_max;
    int64_t* lsum = (int64_t*)lds;
    uint64_t* lmn = (uint64_t*)(lsum + S);
    uint64_t* lmx = lmn + S;
    uint32_t* lcnt = (uint32_t*)(lmx + S);
    int32_t q = W.q_begin;
    int32_t prev_q = -1;
    uint32_t prev_chunks = 0;
    while (true) {
        if (tid == 0) {
            if (prev_q >= 0) mark_done(p, prev_q, prev_chunks);
            const int64_t c = W.chunk_begin + (int64_t)atomicAdd(p.ctr + kCtrStride * wave + kCtrPass2, 1ull);
            s_unit = c;
            if (c < W.chunk_end) {
                while (q < W.q_end && (p.q[q].chunks == 0 || p.q[q].chunk0 + p.q[q].chunks <= c)) q++;
                s_q = q;
            }
        }
        __syncthreads();
        const int64_t c = uniform64(s_unit);
        if (c >= W.chunk_end) break;
        q = (int32_t)uniform64(s_q);
        const BQuery Q = p.q[q];
        const int j = (int)(c - Q.chunk0);
        const int nch = Q.chunks;
        const int32_t key0 = j << Q.cbits;
        const int32_t width = min(1 << Q.cbits, Q.mod - key0);
        for (int g = tid; g < width; g += bd) {
            lsum[g] = 0;
            lmn[g] = ~0ull;
            lmx[g] = 0;
            lcnt[g] = 0;
        }
        __syncthreads();
        // wavefront w takes units w, w + nwv, ...; its lanes stride over the unit's segment
        for (int k = wid; k < Q.units; k += nwv) {
            const int64_t base = p.segb[Q.segb + k];
            if (base < 0) continue;  // that unit's rows took the atomic path
            const int32_t* d = p.desc + Q.desc + (int64_t)k * (nch + 1);
            const int32_t lo = d[j], hi = d[j + 1];
            constexpr int kU = 4;  // rows per lane per iteration, loaded before any is added
            for (int32_t i0 = lo; i0 < hi; i0 += 64 * kU) {
                uint32_t kk[kU];
                int64_t vv[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int32_t i = min(i0 + u * 64 + lane, hi - 1);
                    kk[u] = p.sk[base + i];
                    vv[u] = p.sv[base + i];
                }
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    if (i0 + u * 64 + lane >= hi) continue;
                    const uint32_t g = kk[u] - (uint32_t)key0;
                    const uint64_t e = enc_i64(vv[u]);
                    atomicAdd(&lcnt[g], 1u);
                    atomicAdd((unsigned long long*)&lsum[g], (unsigned long long)vv[u]);
                    atomicMin((unsigned long long*)&lmn[g], (unsigned long long)e);
                    atomicMax((unsigned long long*)&lmx[g], (unsigned long long)e);
                }
            }
        }
        __syncthreads();
        // this workgroup is the only writer of these groups in this kernel
        for (int g = tid; g < width; g += bd) {
            if (!lcnt[g]) continue;
            const int64_t at = Q.out + key0 + g;
            p.cnt[at] += lcnt[g];
            p.sum[at] += (unsigned long long)lsum[g];
            if (lmn[g] < p.mn[at]) p.mn[at] = lmn[g];
            if (lmx[g] > p.mx[at]) p.mx[at] = lmx[g];
        }
        __syncthreads();
        prev_q = q;
        prev_chunks = (uint32_t)nch;
    }
    if (tid == 0 && prev_q >= 0) mark_done(p, prev_q, prev_chunks);
}

__global__ void __launch_bounds__(256) k_key_column(const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                    int32_t b_range, BQuery q, int64_t n, int64_t* out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = key_of(a[i], b[i], b_range, q);
}

uint64_t magic_of(int32_t mod) { return ~0ull / (uint64_t)(uint32_t)mod + 1; }

// Widest key chunk of the two-pass path: the largest power of two <= max(lds_groups, 1024).
int chunk_bits_of(int32_t lds_groups) {
    int b = 10;
    while (b < kMaxKeyBits && (1 << (b + 1)) <= lds_groups) b++;
    return b;
}

// Chunk width of one large query: the widest chunk is narrowed while a chunk would hold more
// than `target` of the query's estimated rows, down to 64 groups, with at most
// 2^(24 - widest bits) chunks (the pass-1 histogram).
int chunk_bits_for(int32_t mod, int64_t est, int cbmax, int64_t target) {
    const int64_t max_chunks = (int64_t)1 << (kMaxKeyBits - cbmax);
    int cb = cbmax;
    while (cb > kMinChunkBits) {
        if ((double)est * (double)((int64_t)1 << cb) <= (double)target * (double)mod) break;
        const int nb = cb - 1;
        if ((((int64_t)mod + (1 << nb) - 1) >> nb) > max_chunks) break;
        cb = nb;
    }
    return cb;
}

size_t pass1_lds_bytes(const qk_batch_tuning& t) {
    const int cb = chunk_bits_of(t.lds_groups);
    const size_t small = (size_t)t.lds_groups * 28;
    // pass bits of a tile, then hist + cursors for 2^24 keys
    const size_t large = (size_t)t.tile_rows / 8 + ((size_t)8 << (kMaxKeyBits - cb));
    return std::max(small, large) + 64;
}

size_t pass2_lds_bytes(const qk_batch_tuning& t) { return ((size_t)28 << chunk_bits_of(t.lds_groups)) + 64; }

}  // namespace
}  // namespace qk

using namespace qk;

struct qk_batch_s {
    qk_batch_columns cols{};
    qk_batch_tuning tune{};
    int64_t scratch_rows = kDefaultScratchRows;
    // device buffers (grown as needed)
    void* plan = nullptr;  // BHeader, BWave[wave_cap], BQuery[q_cap]
    uint32_t* done = nullptr;
    unsigned long long* comp = nullptr;
    int32_t q_cap = 0;
    unsigned long long* ctr = nullptr;  // [0] start clock, then 3 per wave
    int32_t wave_cap = 0;
    BUnit* units = nullptr;
    int64_t unit_cap = 0;
    int32_t* desc = nullptr;
    int64_t desc_cap = 0;
    int64_t* segb = nullptr;
    int64_t segb_cap = 0;
    uint32_t* sk = nullptr;
    int64_t* sv = nullptr;
    int64_t sk_cap = 0;
    unsigned long long *cnt = nullptr, *sum = nullptr, *mn = nullptr, *mx = nullptr;  // owned results
    int64_t out_cap = 0;
    unsigned long long* res[4] = {};  // result arrays of the last run: owned or the caller's
    // pinned copies of the plan and the unit list
    void* host_plan = nullptr;
    BUnit* host_units = nullptr;
    int64_t host_unit_cap = 0;
    std::vector<int64_t> offsets;
    std::vector<int64_t> est_host;
    // sorted ts samples of each block (taken at creation) for the row estimates
    std::vector<int64_t> samples;
    int32_t per_block = 0;
    int32_t last_nq = 0;
    // captured graph of init -> waves -> finalize, valid while the launch arguments match
    hipGraphExec_t graph = nullptr;
    int32_t graph_waves = -1;
    BArgs graph_args{};
};

namespace qk {
namespace {

size_t plan_bytes(int32_t wave_cap, int32_t q_cap) {
    return sizeof(BHeader) + sizeof(BWave) * (size_t)wave_cap + sizeof(BQuery) * (size_t)q_cap;
}

void free_graph(qk_batch_s* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_waves = -1;
}

template <typename T>
int grow(T** p, int64_t* cap, int64_t need, int64_t min_cap) {
    if (*p && *cap >= need) return QK_SUCCESS;
    const int64_t c = std::max(need, std::max(min_cap, *cap + *cap / 2));
    if (*p) pool_free(*p);
    *p = nullptr;
    *cap = 0;
    QK_TRY(pool_alloc((void**)p, (size_t)c * sizeof(T)));
    *cap = c;
    return QK_SUCCESS;
}

void free_plan(qk_batch_s* b) {
    void* ptrs[] = {b->plan, b->done, b->comp, b->ctr};
    for (void* p : ptrs)
        if (p) pool_free(p);
    b->plan = nullptr;
    b->done = nullptr;
    b->comp = nullptr;
    b->ctr = nullptr;
    if (b->host_plan) (void)hipHostFree(b->host_plan);
    b->host_plan = nullptr;
    b->q_cap = b->wave_cap = 0;
}

void free_buffers(qk_batch_s* b) {
    free_plan(b);
    void* ptrs[] = {b->units, b->desc, b->segb, b->sk, b->sv, b->cnt, b->sum, b->mn, b->mx};
    for (void* p : ptrs)
        if (p) pool_free(p);
    b->units = nullptr;
    b->desc = nullptr;
    b->segb = nullptr;
    b->sk = nullptr;
    b->sv = nullptr;
    b->cnt = b->sum = b->mn = b->mx = nullptr;
    for (auto& r : b->res) r = nullptr;
    b->unit_cap = b->desc_cap = b->segb_cap = b->sk_cap = b->out_cap = 0;
    if (b->host_units) (void)hipHostFree(b->host_units);
    b->host_units = nullptr;
    b->host_unit_cap = 0;
    free_graph(b);
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
    p.chunk_max = 1 << chunk_bits_of(b->tune.lds_groups);
    p.hdr = (const BHeader*)b->plan;
    p.waves = (const BWave*)((const char*)b->plan + sizeof(BHeader));
    p.q = (const BQuery*)((const char*)b->plan + sizeof(BHeader) + sizeof(BWave) * (size_t)b->wave_cap);
    p.units = b->units;
    p.t0 = b->ctr;
    p.ctr = b->ctr + 1;
    p.done = b->done;
    p.comp = b->comp;
    p.desc = b->desc;
    p.segb = b->segb;
    p.sk = b->sk;
    p.sv = b->sv;
    p.cnt = b->res[0];
    p.sum = b->res[1];
    p.mn = b->res[2];
    p.mx = b->res[3];
    return p;
}

int launch_all(qk_batch_s* b, int32_t nwaves, hipStream_t s) {
    const BArgs p = args_of(b);
    const int cus = device_info().cus;
    k_batch_init<<<cus * 8, 256, 0, s>>>(p);
    QK_HIP(hipGetLastError());
    for (int32_t w = 0; w < nwaves; w++) {
        k_batch_pass1<<<cus * b->tune.wgs_per_cu, kBThreads, pass1_lds_bytes(b->tune), s>>>(p, w);
        QK_HIP(hipGetLastError());
        k_batch_pass2<<<cus * b->tune.wgs_per_cu, kBThreads, pass2_lds_bytes(b->tune), s>>>(p, w);
        QK_HIP(hipGetLastError());
    }
    k_batch_finalize<<<cus * 8, 256, 0, s>>>(p);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

// Capture init -> waves -> finalize once for the current buffers, tuning and wave count.
void build_graph(qk_batch_s* b, int32_t nwaves) {
    free_graph(b);
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        const int lst = launch_all(b, nwaves, cs);
        const hipError_t e = hipStreamEndCapture(cs, &g);
        ok = lst == QK_SUCCESS && e == hipSuccess;
    }
    if (ok && hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0) != hipSuccess) {
        b->graph = nullptr;
        ok = false;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    (void)hipGetLastError();
    b->graph_waves = ok ? nwaves : -1;  // without a graph the kernels are launched one by one
}

int check_tuning(const qk_batch_tuning& t) {
    if (t.tile_rows < 4096 || t.tile_rows % 4096) return fail(QK_ERROR, "qk_batch: tile_rows must be a positive multiple of 4096");
    const size_t limit = std::min<size_t>(device_info().lds_per_block, 160 * 1024) - 512;
    if (t.lds_groups < 0 || pass1_lds_bytes(t) > limit || pass2_lds_bytes(t) > limit)
        return fail(QK_ERROR, "qk_batch: tile_rows and lds_groups do not fit the LDS");
    if (t.wgs_per_cu < 1 || t.wgs_per_cu > 2) return fail(QK_ERROR, "qk_batch: wgs_per_cu must be 1 or 2");
    const size_t half = std::min<size_t>(device_info().lds_per_cu, 160 * 1024) / 2 - 512;
    if (t.wgs_per_cu == 2 && (pass1_lds_bytes(t) > half || pass2_lds_bytes(t) > half))
        return fail(QK_ERROR, "qk_batch: two workgroups per CU need tile_rows and lds_groups below half the LDS");
    return QK_SUCCESS;
}

// Plan buffers for nq queries in nwaves waves (capacities only grow; the graph keeps pointers).
int ensure_plan(qk_batch_s* b, int32_t nq, int32_t nwaves) {
    if (b->plan && nq <= b->q_cap && nwaves <= b->wave_cap) return QK_SUCCESS;
    free_graph(b);
    const int32_t qc = std::max(nq, std::max(b->q_cap * 2, 1024));
    const int32_t wc = std::max(nwaves, std::max(b->wave_cap * 2, 64));
    free_plan(b);
    const size_t pc = plan_bytes(wc, qc);
    QK_TRY(pool_alloc(&b->plan, pc));
    QK_TRY(pool_alloc((void**)&b->done, (size_t)qc * 4));
    QK_TRY(pool_alloc((void**)&b->comp, (size_t)qc * 8));
    QK_TRY(pool_alloc((void**)&b->ctr, (size_t)(1 + kCtrStride * wc) * 8));
    if (hipHostMalloc(&b->host_plan, pc) != hipSuccess) {
        b->host_plan = nullptr;
        return fail(QK_OUT_OF_MEMORY, "qk_batch: pinned plan buffer");
    }
    b->q_cap = qc;
    b->wave_cap = wc;
    return QK_SUCCESS;
}

int ensure_units(qk_batch_s* b, int64_t n) {
    if (b->units && b->host_units && n <= b->unit_cap && n <= b->host_unit_cap) return QK_SUCCESS;
    free_graph(b);
    const int64_t c = std::max(n, std::max(b->unit_cap * 2, (int64_t)1 << 16));
    if (b->units) pool_free(b->units);
    b->units = nullptr;
    b->unit_cap = 0;
    if (b->host_units) (void)hipHostFree(b->host_units);
    b->host_units = nullptr;
    b->host_unit_cap = 0;
    QK_TRY(pool_alloc((void**)&b->units, (size_t)c * sizeof(BUnit)));
    if (hipHostMalloc((void**)&b->host_units, (size_t)c * sizeof(BUnit)) != hipSuccess) {
        b->host_units = nullptr;
        return fail(QK_OUT_OF_MEMORY, "qk_batch: pinned unit buffer");
    }
    b->unit_cap = b->host_unit_cap = c;
    return QK_SUCCESS;
}

// Units of queries [q_begin, q_end) of the plan, appended at out[*n] as kXcds lists: list x
// holds the tiles t with t mod kXcds == x, tile by tile (the queries of a tile in plan order).
void list_units(const std::vector<BQuery>& bq, BWave* w, BUnit* out, int64_t* n) {
    int64_t t_lo = INT64_MAX, t_hi = 0;
    for (int32_t k = w->q_begin; k < w->q_end; k++) {
        t_lo = std::min(t_lo, bq[(size_t)k].tile0);
        t_hi = std::max(t_hi, bq[(size_t)k].tile0 + bq[(size_t)k].units);
    }
    if (w->q_begin >= w->q_end) t_lo = t_hi = 0;
    std::vector<int64_t> at((size_t)(t_hi - t_lo + 1), 0);
    for (int32_t k = w->q_begin; k < w->q_end; k++) {
        at[(size_t)(bq[(size_t)k].tile0 - t_lo)] += 1;
        at[(size_t)(bq[(size_t)k].tile0 + bq[(size_t)k].units - t_lo)] -= 1;
    }
    std::vector<int64_t> cover((size_t)(t_hi - t_lo), 0);  // queries covering each tile
    int64_t run = 0;
    for (size_t i = 0; i < cover.size(); i++) cover[i] = run += at[i];
    int64_t pos = *n;
    for (int x = 0; x < kXcds; x++) {
        w->unit_off[x] = pos;
        int64_t t = t_lo + (((int64_t)x - t_lo) % kXcds + kXcds) % kXcds;  // first tile of list x
        for (; t < t_hi; t += kXcds) {
            at[(size_t)(t - t_lo)] = pos;
            pos += cover[(size_t)(t - t_lo)];
        }
    }
    w->unit_off[kXcds] = pos;
    for (int32_t k = w->q_begin; k < w->q_end; k++) {
        const BQuery& d = bq[(size_t)k];
        for (int64_t t = d.tile0; t < d.tile0 + d.units; t++) out[at[(size_t)(t - t_lo)]++] = BUnit{k, (int32_t)t};
    }
    *n = pos;
}

// Sorted ts samples of every block, taken once (the table is device-resident and read-only
// for the engine; qk_batch_resample takes them again after the caller changes it).
int take_samples(qk_batch_s* b) {
    const qk_batch_columns& c = b->cols;
    const int64_t nblocks = (c.rows + c.block_rows - 1) / c.block_rows;
    const int32_t per = (int32_t)std::min<int64_t>(kSamples, c.block_rows);
    b->samples.assign((size_t)(nblocks * per), 0);
    b->per_block = per;
    if (nblocks == 0) return QK_SUCCESS;
    int64_t* dev = nullptr;
    QK_TRY(pool_alloc((void**)&dev, (size_t)(nblocks * per) * 8));
    const int grid = (int)std::min<int64_t>((nblocks * per + 255) / 256, 4096);
    k_batch_sample_ts<<<grid, 256>>>(c.ts, c.rows, c.block_rows, per, nblocks, dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(b->samples.data(), dev, (size_t)(nblocks * per) * 8, hipMemcpyDeviceToHost);
    pool_free(dev);
    if (e != hipSuccess) return fail(QK_RUNTIME_ERROR, "qk_batch: sampling ts: %s", hipGetErrorString(e));
    for (int64_t blk = 0; blk < nblocks; blk++)
        std::sort(b->samples.begin() + blk * per, b->samples.begin() + (blk + 1) * per);
    return QK_SUCCESS;
}

// Passing rows of a query estimated from the samples of its blocks.
int64_t estimate_rows(const qk_batch_s* b, const qk_batch_query& x) {
    const qk_batch_columns& c = b->cols;
    const int32_t per = b->per_block;
    if (per == 0 || x.ts_hi <= x.ts_lo) return 0;
    double rows = 0.0;
    for (int64_t blk = x.first_block; blk < (int64_t)x.first_block + x.nblocks; blk++) {
        const int64_t* s0 = b->samples.data() + blk * per;
        const int64_t hit = std::lower_bound(s0, s0 + per, x.ts_hi) - std::lower_bound(s0, s0 + per, x.ts_lo);
        const int64_t len = std::min<int64_t>(c.block_rows, c.rows - blk * c.block_rows);
        rows += (double)hit * (double)len / per;
    }
    return (int64_t)rows;
}

// ext: four caller arrays (count, sum, min, max) of at least sum(key_mod) entries to write the
// results to instead of the engine's own arrays, or null.
int run_batch(qk_batch_s* b, int32_t nq, const qk_batch_query* q, hipStream_t s, unsigned long long* const* ext) {
    const qk_batch_columns& c = b->cols;
    const qk_batch_tuning& t = b->tune;
    const int64_t nblocks = (c.rows + c.block_rows - 1) / c.block_rows;
    const int64_t T = t.tile_rows;
    const int cbmax = chunk_bits_of(t.lds_groups);
    if (nq > (1 << 24)) return fail(QK_ERROR, "qk_batch: at most 2^24 queries per run");
    if ((c.rows + T - 1) / T > INT32_MAX) return fail(QK_ERROR, "qk_batch: too many tiles; raise tile_rows");
    // the previous run of this engine may still read the buffers and pinned plan changed below
    QK_HIP(hipStreamSynchronize(s));
    // queries in plan order: small ones first, then large ones (waves are cut below)
    std::vector<int32_t> order;
    order.reserve((size_t)nq);
    for (int pass = 0; pass < 2; pass++)
        for (int32_t i = 0; i < nq; i++)
            if ((q[i].key_mod > t.lds_groups) == (pass == 1)) order.push_back(i);
    b->offsets.assign((size_t)nq + 1, 0);
    int64_t out = 0;
    for (int32_t i = 0; i < nq; i++) {
        const qk_batch_query& x = q[i];
        if (x.key_mod < 1 || x.key_mod > (1 << kMaxKeyBits))
            return fail(QK_ERROR, "qk_batch: query %d: key_mod must be 1..2^24", i);
        if (x.first_block < 0 || x.nblocks < 1 || x.first_block + (int64_t)x.nblocks > nblocks)
            return fail(QK_ERROR, "qk_batch: query %d: block range outside the table", i);
        b->offsets[(size_t)i] = out;
        out += x.key_mod;
    }
    b->offsets[(size_t)nq] = out;
    std::vector<BQuery> bq((size_t)nq);
    int32_t nlarge = 0;
    for (size_t k = 0; k < order.size(); k++) {
        const int32_t i = order[k];
        const qk_batch_query& x = q[i];
        BQuery& d = bq[k];
        d = BQuery{};
        d.row0 = (int64_t)x.first_block * c.block_rows;
        d.rows = std::min<int64_t>((int64_t)(x.first_block + x.nblocks) * c.block_rows, c.rows) - d.row0;
        d.tile0 = d.row0 / T;
        d.units = (int32_t)((d.row0 + d.rows + T - 1) / T - d.tile0);
        d.ts_lo = x.ts_lo;
        d.ts_hi = x.ts_hi;
        d.out = b->offsets[(size_t)i];
        d.mod = x.key_mod;
        d.magic = magic_of(x.key_mod);
        d.index = i;
        nlarge += x.key_mod > t.lds_groups;
    }
    const int32_t nsmall = nq - nlarge;
    // grow the buffers that depend on the batch
    if (ext) {
        for (int i = 0; i < 4; i++) b->res[i] = ext[i];
    } else if (out > b->out_cap) {
        free_graph(b);
        unsigned long long** arrs[] = {&b->cnt, &b->sum, &b->mn, &b->mx};
        const int64_t cap = std::max(out, std::max(b->out_cap + b->out_cap / 2, (int64_t)1 << 20));
        for (unsigned long long** a : arrs) {
            if (*a) pool_free(*a);
            *a = nullptr;
        }
        b->out_cap = 0;
        for (unsigned long long** a : arrs) QK_TRY(pool_alloc((void**)a, (size_t)cap * 8));
        b->out_cap = cap;
    }
    if (!ext) {
        b->res[0] = b->cnt;
        b->res[1] = b->sum;
        b->res[2] = b->mn;
        b->res[3] = b->mx;
    }
    QK_TRY(ensure_plan(b, nq, std::max(1, nlarge)));
    // estimated passing rows of the large queries, to size the pool and chunks and cut waves
    b->est_host.resize((size_t)nlarge);
    for (int32_t k = nsmall; k < nq; k++) b->est_host[(size_t)(k - nsmall)] = estimate_rows(b, q[bq[(size_t)k].index]);
    // the pool: the batch's estimated rows up to scratch_rows (grown, never shrunk)
    int64_t need_all = 0;
    for (int64_t e : b->est_host) need_all += e + e / 4 + 4096;
    const int64_t want = std::min(b->scratch_rows, need_all);
    if (nlarge && (!b->sk || !b->sv || b->sk_cap < want)) {
        free_graph(b);
        const int64_t cap = std::min(b->scratch_rows, std::max(want, b->sk_cap + b->sk_cap / 2));
        if (b->sk) pool_free(b->sk);
        if (b->sv) pool_free(b->sv);
        b->sk = nullptr;
        b->sv = nullptr;
        b->sk_cap = 0;
        QK_TRY(pool_alloc((void**)&b->sk, (size_t)cap * 4));
        QK_TRY(pool_alloc((void**)&b->sv, (size_t)cap * 8));
        b->sk_cap = cap;
    }
    const int64_t pool = b->sk_cap;
    // chunks of the large queries
    const int64_t target = std::max<int64_t>((int64_t)1 << 18, pool / ((int64_t)device_info().cus * 4));
    int64_t desc = 0, segb = 0, chunk_units = 0;
    for (int32_t k = nsmall; k < nq; k++) {
        BQuery& d = bq[(size_t)k];
        d.cbits = chunk_bits_for(d.mod, b->est_host[(size_t)(k - nsmall)], cbmax, target);
        d.chunks = (int32_t)(((int64_t)d.mod + (1 << d.cbits) - 1) >> d.cbits);
        d.desc = desc;
        d.segb = segb;
        d.chunk0 = chunk_units;
        desc += (int64_t)d.units * (d.chunks + 1);
        segb += d.units;
        chunk_units += d.chunks;
    }
    QK_TRY(grow(&b->desc, &b->desc_cap, std::max<int64_t>(desc, 1), (int64_t)1 << 16));
    QK_TRY(grow(&b->segb, &b->segb_cap, std::max<int64_t>(segb, 1), (int64_t)1 << 12));
    // waves: wave 0 holds every small query; large queries are added while their estimated
    // rows fit the pool (a query above the pool gets a wave of its own)
    std::vector<BWave> waves;
    BWave w{};
    int64_t used = 0;
    w.q_end = nsmall;
    for (int32_t k = nsmall; k < nq; k++) {
        const BQuery& d = bq[(size_t)k];
        const int64_t e = b->est_host[(size_t)(k - nsmall)];
        const int64_t need = e + e / 4 + 4096;
        if (used > 0 && used + need > pool) {
            waves.push_back(w);
            w = BWave{};
            w.q_begin = k;
            w.chunk_begin = d.chunk0;
            used = 0;
        }
        used += need;
        w.q_end = k + 1;
        w.chunk_end = d.chunk0 + d.chunks;
    }
    waves.push_back(w);
    const int32_t nw = (int32_t)waves.size();
    int64_t nunits = 0;
    for (const BQuery& d : bq) nunits += d.units;
    QK_TRY(ensure_units(b, std::max<int64_t>(nunits, 1)));
    QK_TRY(ensure_plan(b, nq, nw));
    int64_t pos = 0;
    for (BWave& x : waves) list_units(bq, &x, b->host_units, &pos);
    BHeader hd{};
    hd.nq = nq;
    hd.nwaves = nw;
    hd.out_total = out;
    hd.scratch_cap = nlarge ? pool : 0;
    hd.nunits = nunits;
    BQuery* host_q = (BQuery*)((char*)b->host_plan + sizeof(BHeader) + sizeof(BWave) * (size_t)b->wave_cap);
    memcpy(b->host_plan, &hd, sizeof hd);
    memcpy((char*)b->host_plan + sizeof hd, waves.data(), sizeof(BWave) * (size_t)nw);
    if (nq) memcpy(host_q, bq.data(), sizeof(BQuery) * (size_t)nq);
    const BArgs p = args_of(b);
    QK_HIP(hipMemcpyAsync(b->plan, b->host_plan, sizeof hd + sizeof(BWave) * (size_t)nw, hipMemcpyHostToDevice, s));
    if (nq) QK_HIP(hipMemcpyAsync((void*)p.q, host_q, sizeof(BQuery) * (size_t)nq, hipMemcpyHostToDevice, s));
    if (nunits) QK_HIP(hipMemcpyAsync(b->units, b->host_units, sizeof(BUnit) * (size_t)nunits, hipMemcpyHostToDevice, s));
    b->last_nq = nq;
    if (t.use_graph) {
        if (b->graph_waves != nw || memcmp(&p, &b->graph_args, sizeof p) != 0) {
            build_graph(b, nw);
            b->graph_args = p;
        }
        if (b->graph_waves == nw) {
            QK_HIP(hipGraphLaunch(b->graph, s));
            return QK_SUCCESS;
        }
    }
    return launch_all(b, nw, s);
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
    b->tune.scratch_rows = 0;
    const int st = take_samples(b);
    if (st != QK_SUCCESS) {
        delete b;
        return st;
    }
    *out = b;
    return QK_SUCCESS;
}

int qk_batch_resample(qk_batch b) {
    if (!b) return fail(QK_ERROR, "qk_batch_resample: null batch");
    QK_TRY(ensure_init());
    return take_samples(b);
}

int qk_batch_estimate_rows(qk_batch b, int32_t nq, const qk_batch_query* q, int64_t* out) {
    if (!b || nq < 0 || (nq > 0 && (!q || !out))) return fail(QK_ERROR, "qk_batch_estimate_rows: bad argument");
    const int64_t nblocks = (b->cols.rows + b->cols.block_rows - 1) / b->cols.block_rows;
    for (int32_t i = 0; i < nq; i++) {
        if (q[i].first_block < 0 || q[i].nblocks < 1 || q[i].first_block + (int64_t)q[i].nblocks > nblocks)
            return fail(QK_ERROR, "qk_batch_estimate_rows: query %d: block range outside the table", i);
        out[i] = estimate_rows(b, q[i]);
    }
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
    if (t->scratch_rows < 0) return fail(QK_ERROR, "qk_batch_set_tuning: scratch_rows must be >= 0");
    const bool same = memcmp(&b->tune, t, sizeof *t) == 0;
    b->tune = *t;
    const int64_t rows = t->scratch_rows > 0 ? t->scratch_rows : kDefaultScratchRows;
    if (rows != b->scratch_rows) {
        if (b->sk) pool_free(b->sk);
        if (b->sv) pool_free(b->sv);
        b->sk = nullptr;
        b->sv = nullptr;
        b->sk_cap = 0;
        b->scratch_rows = rows;
        free_graph(b);
    }
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
    return run_batch(b, nq, q, as_stream(stream), nullptr);
}

int qk_batch_run_into(qk_batch b, int32_t nq, const qk_batch_query* q, int64_t* count, int64_t* sum,
                      int64_t* min, int64_t* max, void* stream) {
    if (!b || (nq > 0 && !q) || nq < 0) return fail(QK_ERROR, "qk_batch_run_into: bad argument");
    if (nq > 0 && (!count || !sum || !min || !max)) return fail(QK_ERROR, "qk_batch_run_into: null result array");
    QK_TRY(ensure_init());
    unsigned long long* ext[4] = {(unsigned long long*)count, (unsigned long long*)sum, (unsigned long long*)min,
                                  (unsigned long long*)max};
    return run_batch(b, nq, q, as_stream(stream), ext);
}

int qk_batch_autotune(qk_batch b, int32_t nq, const qk_batch_query* q, qk_batch_tuning* out, void* stream) {
    if (!b || nq < 1 || !q) return fail(QK_ERROR, "qk_batch_autotune: bad argument");
    QK_TRY(ensure_init());
    hipStream_t s = as_stream(stream);
    const qk_batch_tuning keep = b->tune;
    std::vector<qk_batch_tuning> cands;
    for (int tile : {1 << 16, 1 << 17, 1 << 18, 1 << 19})
        for (int groups : {2048, 4096, 5120})
            for (int wpc : {1, 2}) {
                qk_batch_tuning t{tile, groups, wpc, keep.use_graph, keep.scratch_rows};
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
        st = run_batch(b, nq, q, s, nullptr);  // warm-up (and graph capture)
        if (st != QK_SUCCESS) break;
        if (hipEventRecord(e0, s) != hipSuccess) { st = QK_RUNTIME_ERROR; break; }
        st = run_batch(b, nq, q, s, nullptr);
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
    if (count) *count = (const int64_t*)b->res[0];
    if (sum) *sum = (const int64_t*)b->res[1];
    if (min) *min = (const int64_t*)b->res[2];
    if (max) *max = (const int64_t*)b->res[3];
    if (offsets_host) *offsets_host = b->offsets.data();
    return QK_SUCCESS;
}

int qk_batch_completion_ms(qk_batch b, double* out, int32_t nq) {
    if (!b || !out || nq < 0 || nq > b->last_nq) return fail(QK_ERROR, "qk_batch_completion_ms: bad argument");
    QK_HIP(hipDeviceSynchronize());
    std::vector<unsigned long long> comp((size_t)nq);
    unsigned long long t0 = 0;
    if (nq) QK_HIP(hipMemcpy(comp.data(), b->comp, (size_t)nq * 8, hipMemcpyDeviceToHost));
    QK_HIP(hipMemcpy(&t0, b->ctr, 8, hipMemcpyDeviceToHost));
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
