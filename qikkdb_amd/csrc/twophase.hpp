// twophase.hpp - host interface of the two-phase partitioned GROUP BY (twophase.hip).
//
// Phase 1 streams a block once: rows passing the WHERE whose key is one of the table's heavy
// hitters (sampled from the first block) are aggregated into the workgroup's LDS table and
// written once per workgroup as a partial; every other row is appended, without atomics, to
// the workgroup's private region of its hash partition.  Phase 2 gives each partition to one
// workgroup, which folds the partition's rows and the heavy partials in LDS and merges them
// into the partition's sub-table of the global table, which no other workgroup touches, so
// the merge needs no atomics.  Rows that do not fit (NULL values, full regions) take the
// atomic path in phase 1.
//
// Phase 2 may be deferred over several blocks: phase 1 then appends to the regions and adds
// to the heavy partials, and one phase 2 merges all of them, so the global table is walked
// once per query instead of once per block.  Whatever reads the table flushes first.
#pragma once

#include "groupby.hpp"

namespace qk {

struct TwoPhaseBufs {
    int32_t W = 0;          // phase-1 workgroups
    int32_t P = 1;          // partitions (= sub-tables of the global table)
    int32_t cap = 0;        // rows per (workgroup, partition) region
    int32_t vbytes = 0;     // bytes per value in the regions
    int32_t kmax = 0;       // heavy keys at most
    int32_t H1 = 0;         // phase-1 LDS hash slots
    int32_t H2 = 0;         // phase-2 LDS hash slots
    int32_t block = 1024;   // threads per workgroup (both phases)
    int32_t sample_bits = 18;
    int64_t rows_alloc = 0;   // rows the regions were sized for (several blocks when deferring)
    int64_t block_alloc = 0;  // largest block the buffers were sized for
    int64_t defer_rows = -1;  // rows phase 2 may be deferred over: -1 default, 0 never
    int64_t pending_rows = 0; // rows phase 1 processed since the last phase 2
    int32_t pend_vtype = 0;   // value type of the pending rows (0 = no value column)
    // heavy set
    uint64_t* heavy_keys = nullptr;   // [kmax], sorted by partition
    uint32_t* heavy_cnt = nullptr;    // [kmax] sample count of each heavy key
    int32_t* heavy_meta = nullptr;    // [0] = K, [1] = threshold, [2..] = start per partition (P+1)
    uint64_t* sample_keys = nullptr;  // [2^(sample_bits+1)]
    uint32_t* sample_cnt = nullptr;
    uint32_t* hist = nullptr;         // [kHistBins]
    uint64_t* cand = nullptr;         // [kmax] unsorted heavy candidates
    uint32_t* cand_cnt = nullptr;     // [kmax] their sample counts
    uint64_t* p1_image = nullptr;     // [H1] keys of the phase-1 heavy table, built once per heavy set
    int32_t* p1_slot = nullptr;       // [kmax] slot of each heavy key in it (-1: not placed)
    // per phase-1 workgroup partials [W][kmax + 2]
    uint32_t* p_cnt = nullptr;
    uint32_t* p_rows = nullptr;
    uint64_t* p_sum = nullptr;
    uint64_t* p_min = nullptr;
    uint64_t* p_max = nullptr;
    // cold regions [W][P][cap]
    uint32_t* r_rec = nullptr;        // key + value records
    uint32_t* r_count = nullptr;      // [W][P]
    uint64_t* zero_word = nullptr;    // 64 zero bytes
    GView* gview = nullptr;           // device copy of the table view (phase 1 reads it there)
    GView gview_host{};               // what gview holds
    bool gview_valid = false;
    bool heavy_ready = false;
    int32_t last_form = 0;            // phase-1 kernel of the last block: 1 streaming, 0 general
};

constexpr int kHistBins = 1024;

// Bytes per slot of a sub-table held in LDS by phase 2 (the key, 32-bit count deltas and
// 64-bit state deltas), and the LDS bytes of a sub-table of S slots.  Phase 2 folds straight
// into the global sub-tables when these fit the LDS.
__host__ __device__ inline int owned_slot_bytes(uint32_t fields) {
    return 8 + 4 + ((fields & F_CNTALL) ? 4 : 0) + ((fields & F_SUM) ? 8 : 0) + ((fields & F_MIN) ? 8 : 0) +
           ((fields & F_MAX) ? 8 : 0);
}
__host__ __device__ inline size_t p2o_lds_bytes(int64_t S, uint32_t fields) {
    return (size_t)S * (size_t)owned_slot_bytes(fields);
}

// Phase 1 writes cold rows to at most kMaxRecParts partitions (the top bits of the key hash).
// A table may have up to 2^kMaxSplitBits sub-tables per partition: phase 2 then runs one
// workgroup per sub-table, each reading its partition's rows and keeping its own.
constexpr int32_t kMaxRecParts = 1024;
constexpr int32_t kMaxSplitBits = 4;

// Whether the two-phase path applies: single-word keys and a value type it handles.
bool twophase_supported(int32_t kw, int32_t value_type);

// Whether the buffers must be reallocated for blocks of n rows (which drops pending rows).
bool twophase_needs_realloc(const TwoPhaseBufs& b, int32_t parts, int32_t value_type,
                            bool has_values, int64_t n, bool defer);

// (Re)size the buffers for blocks of n rows; fields/value type decide the LDS layouts.
int twophase_prepare(TwoPhaseBufs& b, int32_t parts, uint32_t fields, int32_t value_type,
                     bool has_values, int64_t n, bool defer, hipStream_t s);
void twophase_free(TwoPhaseBufs& b);

// Sample the block and build the heavy-hitter set (device-side, no host round trip).
int twophase_sample(TwoPhaseBufs& b, const GView& g, const KeySpec& ks, const KeyCols& kc,
                    int32_t vtype, const void* vals, const uint64_t* vnull, const int8_t* mask,
                    const PredSet& ps, uint32_t vpred, int64_t n, int mode, hipStream_t s);

// Phase 1 over one block: heavy rows into the partials, cold rows into the regions.
int twophase_phase1(TwoPhaseBufs& b, const GView& g, const KeySpec& ks, const KeyCols& kc,
                    int32_t vtype, const void* vals, const uint64_t* vnull, const int8_t* mask,
                    const PredSet& ps, uint32_t vpred, int64_t n, uint32_t fields, int mode,
                    hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_end);

// Phase 2 over everything phase 1 left since the last flush (nothing to do if none).
int twophase_flush(TwoPhaseBufs& b, const GView& g, uint32_t fields, hipStream_t s);

}  // namespace qk
