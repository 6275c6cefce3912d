/*
 * qikkdb_amd.h - C ABI of the MI355X (gfx950) WHERE -> reconstruct -> GROUP BY path.
 *
 * Every entry point takes plain pointers, sizes and an optional HIP stream handle
 * (`void* stream`, NULL = the device's default stream).  Device buffers are raw device
 * pointers.  Buffers that an entry point allocates for the caller ("Keep" outputs,
 * GROUP BY results) come from qk_malloc and are released by the caller with qk_free,
 * which is the ownership rule of the reference's GPUMemory::alloc/free
 * (qikkDB/QueryEngine/GPUCore/GPUMemory.cuh:89-110, GPUMemory.cpp:15-20).
 *
 * Each entry point returns a status code.  Codes 0..5 carry the meaning of the
 * reference's QueryEngineErrorType (qikkDB/QueryEngine/GPUError.h:17-28); the C++ drop-in
 * wrappers under include/qikkdb/ rethrow them as query_engine_error.  The text of the last
 * failure on the calling thread is returned by qk_last_error_message().
 */
#ifndef QIKKDB_AMD_H
#define QIKKDB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
enum qk_status {
    QK_SUCCESS = 0,           /* GPU_EXTENSION_SUCCESS */
    QK_ERROR = 1,             /* GPU_EXTENSION_ERROR: invalid argument or unsupported form */
    QK_DIVISION_BY_ZERO = 2,  /* GPU_DIVISION_BY_ZERO_ERROR */
    QK_INTEGER_OVERFLOW = 3,  /* GPU_INTEGER_OVERFLOW_ERROR */
    QK_HASH_TABLE_FULL = 4,   /* GPU_HASH_TABLE_FULL */
    QK_UNKNOWN_AGG_FUN = 5,   /* GPU_UNKNOWN_AGG_FUN */
    QK_RUNTIME_ERROR = 64,    /* a HIP runtime call failed (the reference aborts, GPUError.cpp:37) */
    QK_OUT_OF_MEMORY = 65     /* device allocation failed (the reference throws std::out_of_range) */
};

/* ---- column types: numbering of DataType (qikkDB/DataType.h:18-25) --------------------- */
enum qk_type { QK_INT = 8, QK_LONG = 9, QK_FLOAT = 10, QK_DOUBLE = 11, QK_INT8 = 15 };

/* ---- filter operators: the FilterConditions functors (GPUFilterConditions.cuh:9-110) ---- */
enum qk_filter_op {
    QK_GREATER = 0,
    QK_GREATER_EQUAL = 1,
    QK_LESS = 2,
    QK_LESS_EQUAL = 3,
    QK_EQUAL = 4,
    QK_NOT_EQUAL = 5,
    QK_LOGICAL_AND = 6,
    QK_LOGICAL_OR = 7
};

/* ---- operand kinds: column pointer or scalar (MaybeDeref.cuh:3-14) --------------------- */
enum qk_operand { QK_COLUMN = 0, QK_CONST = 1 };

/* ---- aggregation functors (AggregationFunctions.cuh:17-366) ----------------------------
 * QK_AGG_COUNT counts rows whose value is not NULL (all rows when no value null mask is
 * given, which is how the reference evaluates COUNT(*), GpuSqlDispatcherAggregationFunctions.h
 * :388-406).  QK_AGG_COUNT_ALL counts every row whatever the value null mask says, so that a
 * table holding several aggregates can carry COUNT(*) next to a nullable value column.     */
enum qk_agg {
    QK_AGG_MIN = 0,
    QK_AGG_MAX = 1,
    QK_AGG_SUM = 2,
    QK_AGG_AVG = 3,
    QK_AGG_COUNT = 4,
    QK_AGG_NONE = 5,
    QK_AGG_COUNT_ALL = 6
};

/* ======================================================================================
 * Runtime: device binding, memory, streams
 * ====================================================================================== */

/* Bind the calling thread to `device` and initialise the library on it.
 * Replaces Context::getInstance() + bindDeviceToContext (Context.cu:20-100, :151-162). */
int qk_init(int device);
int qk_device_count(int* count);
/* Text of the last failure on the calling thread ("" when none). */
const char* qk_last_error_message(void);

/* Pooled device allocation.  Replaces GPUMemory::alloc / GPUMemory::free (GPUMemory.cuh:89-110,
 * GPUMemory.cpp:15-20) over the reference's CudaMemAllocator (CudaMemAllocator.cu:109-322). */
int qk_malloc(void** ptr, size_t bytes);
/* The block goes back to the pool at once, and the next qk_malloc of its size may return it,
 * to any stream: free a buffer only after the work that reads it has completed (the
 * reference's GPUMemory::free is cudaFree, which waits for the device itself). */
int qk_free(void* ptr);
/* Bytes handed out and bytes held in the pool's free lists. */
int qk_mem_stats(size_t* in_use, size_t* cached);
/* Return every cached block to the driver. */
int qk_mem_release_cached(void);

/* Cap on the device memory the pool holds from the driver (blocks in use plus cached ones;
 * 0 = no cap).  Past it qk_malloc first returns cached blocks to the driver, then fails with
 * QK_OUT_OF_MEMORY.  Replaces the GPUMemoryUsagePercent reservation of CudaMemAllocator
 * (CudaMemAllocator.cu:109-152); qk_mem_set_limit_percent takes the same percentage, of the
 * memory free now plus what the pool already holds.  Per device (the calling thread's). */
int qk_mem_set_limit(size_t bytes);
int qk_mem_set_limit_percent(int32_t percent);
int qk_mem_get_limit(size_t* bytes);
/* Whether qk_malloc(bytes) would succeed now: CudaMemAllocator::CanAllocate
 * (CudaMemAllocator.h:57). */
int qk_mem_can_allocate(size_t bytes, int32_t* out);

/* Copies: GPUMemory::copyHostToDevice / copyDeviceToHost / copyDeviceToDevice
 * (GPUMemory.cuh:113-189).  qk_memcpy_dtoh returns after the bytes are on the host. */
int qk_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int qk_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int qk_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int qk_memset(void* dst, int byte_value, size_t bytes, void* stream);

int qk_stream_create(void** stream);
int qk_stream_destroy(void* stream);
int qk_stream_synchronize(void* stream);
int qk_device_synchronize(void);

/* ---- block cache: GPUMemoryCache (GPUMemoryCache.h:10-141, GPUMemoryCache.cpp:13-117) ----
 * Device blocks named by (database, table.column, block index, load size, load offset), kept
 * in least-recently-used order under max_bytes and the pool's cap.  A miss allocates the block
 * (from qk_malloc; its contents are the caller's to fill), evicting least recently used blocks
 * first; blocks of the current block index whose "database.table.column" is on the lock list
 * are never evicted.  The cache owns its blocks: they stay valid until evicted, cleared or the
 * cache is destroyed, and the device is synchronised before a block is given up. */
typedef struct qk_cache_s* qk_cache;
/* On the calling thread's device (GPUMemoryCache(deviceID, maximumSize)). */
int qk_cache_create(qk_cache* out, size_t max_bytes);
int qk_cache_destroy(qk_cache c);
/* GetColumn<T>: *hit = 1 when the block was cached.  QK_ERROR when bytes exceeds max_bytes
 * ("Tried to cache block larger than maximum cache size"), QK_OUT_OF_MEMORY when nothing is
 * left to evict ("Not enough space left in cache"); the reference throws std::length_error. */
int qk_cache_get_column(qk_cache c, const char* database, const char* table_column, int32_t block_index,
                        size_t bytes, int64_t load_size, int64_t load_offset, void** ptr, size_t* out_bytes,
                        int32_t* hit);
int qk_cache_contains(qk_cache c, const char* database, const char* table_column, int32_t block_index,
                      int64_t load_size, int64_t load_offset, int32_t* out);
/* Evict: the least recently used unlocked block; *evicted = 0 when there is none. */
int qk_cache_evict(qk_cache c, int32_t* evicted);
/* ClearCachedBlock: every cached load of (database, table_column, block_index). */
int qk_cache_clear_block(qk_cache c, const char* database, const char* table_column, int32_t block_index);
int qk_cache_set_current_block(qk_cache c, int64_t block_index);
/* SetLockList: process-wide, like the reference's static list ("database.table.column"). */
int qk_cache_set_lock_list(int32_t n, const char* const* columns);
int qk_cache_stats(qk_cache c, size_t* used, int64_t* entries, int64_t* hits, int64_t* misses, int64_t* evictions);

/* ======================================================================================
 * Filter (WHERE)
 * ====================================================================================== */

/* out_mask[i] = OP(lhs[i] or lhs, rhs[i] or rhs) && !null(i), one int8 0/1 per row.
 * Comparison typing follows the C++ usual arithmetic conversions of (ltype, rtype).  A
 * QK_CONST operand points to a HOST scalar of its type.  null_mask (LSB-first uint64 words,
 * NULL = no nulls) is the OR of the operands' masks, as the dispatcher prepares it.
 * const-const fills the mask.  Replaces GPUBinary<OP,int8_t,L,R>::Binary
 * (GPUBinary.cuh:554-579) over kernel_filter (GPUFilter.cuh:22-44). */
int qk_filter(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
              int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask,
              int8_t* out_mask, void* stream);

/* Same predicate, written as a bit mask: bit (i % 64) of word i / 64 is row i's result.
 * No reference counterpart (the reference only has int8 masks); fed to the bit-mask forms
 * of the GROUP BY below. */
int qk_filter_bits(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
                   int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask,
                   uint64_t* out_bits, void* stream);

/* out[i] = !in[i] && !null(i).  Replaces kernel_operator_not (GPULogic.cuh:17-33). */
int qk_logical_not(const int8_t* in, int32_t n, const uint64_t* null_mask, int8_t* out,
                   void* stream);

/* ======================================================================================
 * Reconstruct (stream compaction)
 * ====================================================================================== */

/* Inclusive (inclusive != 0) or exclusive int32 prefix sum of an int8 mask.  Replaces
 * GPUReconstruct::PrefixSum / PrefixSumExclusive (GPUReconstruct.cuh:572-601). */
int qk_prefix_sum(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive, void* stream);

/* qk_prefix_sum without any allocation: its look-back state lives in a fixed device array of
 * the library (about 2 MB per device, held until the process ends) instead of per-stream
 * scratch, and the call makes no hipMalloc / hipFree.  No counterpart in the reference; it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops).
 * At most 2^30 rows (QK_ERROR beyond).  qk_init sets the array up; without qk_init the first
 * call on a device does, and synchronises the device then.  Calls on one device share the
 * array: calls that overlap in time (other streams, other threads) race. */
int qk_prefix_sum_noalloc(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive,
                          void* stream);

/* Sum of the int8 mask values, written to *out_host.  Replaces GPUReconstruct::Sum
 * (GPUReconstruct.cuh:604-624) on masks. */
int qk_mask_sum(const int8_t* mask, int32_t n, int64_t* out_host, void* stream);

/* Stable compaction of `in` (n elements of elem_size bytes, 1/2/4/8) by mask, with the
 * null bits of in_null carried along.  Allocates *out (and *out_null when both out_null and
 * in_null are given); *out_count is the number of kept rows.  mask == NULL aliases the input
 * (*out = in, no ownership passes); zero kept rows or n <= 0 give *out = NULL.  Replaces
 * GPUReconstruct::reconstructColKeep (GPUReconstruct.cuh:271-355). */
int qk_reconstruct_col_keep(int32_t elem_size, const void* in, const int8_t* mask, int32_t n,
                            const uint64_t* in_null, void** out, int32_t* out_count,
                            uint64_t** out_null, void* stream);

/* Indices (int32 when index_size == 4, int64 when 8) of the rows whose mask is non-zero, in
 * row order; allocates *out.  Replaces GPUReconstruct::GenerateIndexesKeep
 * (GPUReconstruct.cuh:517-566). */
int qk_generate_indexes_keep(const int8_t* mask, int32_t n, int32_t index_size, void** out,
                             int32_t* out_count, void* stream);

/* One byte per row (bit 0 used) -> LSB-first uint64 bit mask of ceil(n/64) words; allocates
 * *out.  Replaces GPUReconstruct::CompressNullMask (GPUReconstruct.cu:756-764). */
int qk_compress_null_mask(const uint8_t* in, int32_t n, uint64_t** out, void* stream);

/* WHERE + reconstruct in one pass (no reference counterpart: the reference runs the filter,
 * then one scan + scatter per column, GpuSqlDispatcherAggregationFunctions.h:422-442 and
 * :532-536).  Rows passing every predicate are copied, in row order, from each of the ncols
 * input columns into caller-provided outputs of capacity n; *out_count receives the kept
 * row count.  Predicates are described by qk_predicate below. */
struct qk_predicate;
int qk_where_compact(int32_t npreds, const struct qk_predicate* preds, int32_t ncols,
                     const int32_t* elem_sizes, const void* const* cols,
                     const uint64_t* const* col_nulls, int32_t n, void* const* outs,
                     uint64_t* const* out_nulls, int32_t* out_count, void* stream);

/* ======================================================================================
 * GROUP BY
 * ====================================================================================== */

/* One comparison of a conjunctive WHERE, evaluated inside the GROUP BY kernel.
 * column OP (rhs_const | rhs_column); a NULL operand makes the row fail (GPUFilter.cuh:28-43). */
typedef struct qk_predicate {
    int32_t op;                   /* QK_GREATER .. QK_NOT_EQUAL */
    int32_t type;                 /* type of `column` */
    const void* column;           /* device pointer, one element per row of the block */
    const uint64_t* null_mask;    /* NULL = column not nullable */
    int32_t rhs_kind;             /* QK_CONST or QK_COLUMN */
    int32_t rhs_type;             /* type of the constant or of rhs_column */
    const void* rhs_column;
    const uint64_t* rhs_null_mask;
    int64_t rhs_int;              /* constant when rhs_type is QK_INT8/QK_INT/QK_LONG */
    double rhs_real;              /* constant when rhs_type is QK_FLOAT/QK_DOUBLE */
} qk_predicate;

typedef struct qk_groupby_s* qk_groupby;

/* qk_gb_create flags */
enum qk_gb_flags {
    /* The caller promises that no key column is ever nullable: narrower key records. */
    QK_GB_KEYS_NOT_NULL = 1u << 0,
    /* ProcessBlock calls do not wait for the kernel; a HASH_TABLE_FULL latched by one of them
     * is returned by qk_gb_check_error / qk_gb_get_results instead. */
    QK_GB_ASYNC = 1u << 1,
    /* Instead of failing with QK_HASH_TABLE_FULL the table grows itself: after each block it
     * checks itself (a host wait, so such a table neither returns early nor defers its merge),
     * and past its key capacity it moves its groups into a table of at least twice the
     * capacity.  Rows whose keys found no slot are aggregated again into the grown table: they
     * are the rows whose key it lacks, so no block restarts and no row counts twice.  The
     * reference grows and restarts the query instead (GpuSqlDispatcher.cpp:378-394).  The
     * largest table holds 2^30 keys; past that QK_HASH_TABLE_FULL is returned.  Keys cannot be
     * pre-seeded into such a table (qk_gb_seed_keys). */
    QK_GB_AUTO_GROW = 1u << 2
};

/* Create a hash table for GROUP BY over nkeys key columns (nkeys == 1: single key,
 * GPUGroupBy<AGG,O,K,V>, GPUGroupBy.cuh:313-362; nkeys > 1: composite key,
 * GPUGroupBy<AGG,O,std::vector<void*>,V>, GPUGroupByMultiKey.cuh:330-420).
 * aggs[naggs] lists the aggregates kept per group (the reference keeps one per table; a
 * table with several fuses them in one pass).  out_types[naggs] may be NULL for the
 * reference typing (COUNT -> QK_LONG, others -> value_type) or widen SUM (QK_INT -> QK_LONG,
 * QK_FLOAT -> QK_DOUBLE).  value_type is ignored when only COUNT/COUNT_ALL/NONE are kept.
 * The table accepts max_hash_count * key_size_multiplier distinct non-NULL keys (capped at
 * 2^30, IGroupBy.h:6) before it reports QK_HASH_TABLE_FULL, like the reference. */
int qk_gb_create(qk_groupby* out, int32_t nkeys, const int32_t* key_types, int32_t value_type,
                 int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                 int32_t max_hash_count, int32_t key_size_multiplier, uint32_t flags);
int qk_gb_destroy(qk_groupby h);
/* Drop every group; the table is then as freshly created. */
int qk_gb_reset(qk_groupby h, void* stream);
/* GPUGroupBy::GetMaxHashCount (GPUGroupBy.cuh:414-417): the key capacity (of a growing table:
 * its current capacity). */
int qk_gb_get_max_hash_count(qk_groupby h, int32_t* out);

/* Aggregate one block of n rows.  keys[nkeys] are device columns; key_nulls is NULL or an
 * array of nkeys masks (each may be NULL); values may be NULL when only COUNT/NONE are kept;
 * value_nulls may be NULL.  Replaces GPUGroupBy::ProcessBlock (GPUGroupBy.cuh:395-410 and
 * GPUGroupByMultiKey.cuh:535-600): unless the table is QK_GB_ASYNC the call waits and returns
 * QK_HASH_TABLE_FULL like ErrorFlagSwapper::Swap (ErrorFlagSwapper.h:32-48). */
int qk_gb_process_block(qk_groupby h, const void* const* keys, const uint64_t* const* key_nulls,
                        const void* values, const uint64_t* value_nulls, int32_t n, void* stream);

/* ProcessBlock restricted to rows whose int8 mask is non-zero: the same groups as compacting
 * keys and values by the mask first (reconstructColKeep) and calling ProcessBlock on the
 * result, which is the dispatcher's sequence (GpuSqlDispatcherAggregationFunctions.h:422-455,
 * :532-536), without materialising the compacted columns. */
int qk_gb_process_block_masked(qk_groupby h, const void* const* keys,
                               const uint64_t* const* key_nulls, const void* values,
                               const uint64_t* value_nulls, const int8_t* mask, int32_t n,
                               void* stream);

/* ProcessBlock restricted to rows that pass every one of npreds predicates (a conjunctive
 * WHERE), evaluated in the same kernel.  Same groups as filter -> AND -> compact ->
 * ProcessBlock in the reference (GpuSqlDispatcherBinaryFunctions.h:9-158,
 * GpuSqlDispatcher.cpp:1107-1114, GpuSqlDispatcherAggregationFunctions.h:377-553). */
int qk_gb_process_block_where(qk_groupby h, const void* const* keys,
                              const uint64_t* const* key_nulls, const void* values,
                              const uint64_t* value_nulls, int32_t n, int32_t npreds,
                              const qk_predicate* preds, void* stream);

/* Wait for the table's work and return its latched error (QK_SUCCESS or QK_HASH_TABLE_FULL),
 * clearing it. */
int qk_gb_check_error(qk_groupby h, void* stream);

/* Final results: one row per group, in hash-slot order (unspecified, as in the reference).
 * out_keys[nkeys] receive device columns of the key types; out_key_nulls (NULL or
 * [nkeys]) receive key null bit masks; out_values[naggs] receive one device column per
 * aggregate (NULL for QK_AGG_NONE); out_value_nulls (NULL or [naggs]) receive value null bit
 * masks (bit set = no non-NULL value reached the group; COUNT never NULL).  All allocated
 * with qk_malloc; zero groups give NULL pointers and *out_count == 0.  Replaces
 * GPUGroupBy::GetResults (GPUGroupBy.cuh:478-587, GPUGroupByMultiKey.cuh:742-923). */
int qk_gb_get_results(qk_groupby h, void** out_keys, uint64_t** out_key_nulls,
                      void** out_values, uint64_t** out_value_nulls, int32_t* out_count,
                      void* stream);

/* As qk_gb_get_results, with other outputs computed from the aggregate states the table keeps:
 * aggs[naggs] may name any aggregate whose state exists (COUNT always; SUM and AVG when the table
 * keeps a sum, i.e. it was created with SUM or AVG; MIN; MAX; COUNT_ALL), out_types as in
 * qk_gb_create (NULL or 0 = reference typing).  This is how the raw numbers of
 * GPUGroupBy::ReconstructRawNumbers (GPUGroupBy.cuh:433-471: merged sums and occurrence counts
 * of an AVG table, etc.) are read. */
int qk_gb_get_results_aggs(qk_groupby h, int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                           void** out_keys, uint64_t** out_key_nulls, void** out_values,
                           uint64_t** out_value_nulls, int32_t* out_count, void* stream);

/* Pre-seed a fresh single-key table with n distinct keys (a device column of the table's key
 * type; key 0 stands for the NULL group when first_key_is_null): each becomes a group even if
 * no row reaches it (COUNT 0, other aggregates NULL), and qk_gb_get_results returns the
 * seeded groups first, in the order of `keys`, then any other group.  Replaces the
 * constructor GPUGroupBy(maxHashCount, K* keys, bool firstKeyIsNull) (GPUGroupBy.cuh:367-372),
 * whose table keeps its keys at consecutive slots so that GetResults follows their order.
 * QK_ERROR when the table already holds groups or the keys are not distinct. */
int qk_gb_seed_keys(qk_groupby h, const void* keys, int32_t n, int32_t first_key_is_null, void* stream);

/* Fold every group of table src into table dst (tables of the same key and aggregate layout on
 * the calling thread's device): afterwards dst holds what one table that saw both tables' rows
 * would hold; src is unchanged.  Replaces the multi-table GetResults(..., tables, ...) merge
 * (GPUGroupBy.cuh:594-790, GPUGroupByMultiKey.cuh:924-1324), which copies every table's
 * results to the host and re-aggregates them there; here the states (count, sum, min, max) are
 * merged on the device, so AVG stays exact.  Waits; QK_HASH_TABLE_FULL when dst fills up. */
int qk_gb_merge(qk_groupby dst, qk_groupby src, void* stream);

/* ---- GROUP BY over one string key column: GPUGroupBy<AGG,O,std::string,V>
 * (GPUGroupByString.cuh:39-785).  A string column is the reference's GPUString
 * (GPUMemory.cuh:70-77): chars = every string's bytes concatenated, indices[i] = end offset of
 * string i (string i is chars[indices[i-1] .. indices[i]), indices[-1] = 0), both device
 * pointers.  Aggregates, value types, NULL rules, capacity and errors are those of qk_gb_*
 * (the table holds max_hash_count * key_size_multiplier distinct non-NULL strings).  Every
 * ProcessBlock waits (it reads the block's byte count and the dictionary's state back). */
typedef struct qk_gbs_s* qk_gbs;
int qk_gbs_create(qk_gbs* out, int32_t value_type, int32_t naggs, const int32_t* aggs, const int32_t* out_types,
                  int32_t max_hash_count, int32_t key_size_multiplier, uint32_t flags);
int qk_gbs_destroy(qk_gbs h);
int qk_gbs_reset(qk_gbs h, void* stream);
int qk_gbs_process_block(qk_gbs h, const char* chars, const int64_t* indices, const uint64_t* key_nulls,
                         const void* values, const uint64_t* value_nulls, int32_t n, void* stream);
/* As qk_gbs_process_block, for the rows that pass every predicate (a conjunctive WHERE). */
int qk_gbs_process_block_where(qk_gbs h, const char* chars, const int64_t* indices, const uint64_t* key_nulls,
                               const void* values, const uint64_t* value_nulls, int32_t n, int32_t npreds,
                               const qk_predicate* preds, void* stream);
/* One row per group: the keys as a GPUString (out_chars, out_indices; the NULL group's string
 * is empty and its bit set in out_key_nulls), the aggregates as in qk_gb_get_results.  All
 * outputs come from qk_malloc. */
int qk_gbs_get_results(qk_gbs h, char** out_chars, int64_t** out_indices, uint64_t** out_key_nulls,
                       void** out_values, uint64_t** out_value_nulls, int32_t* out_count, void* stream);
int qk_gbs_group_count(qk_gbs h, int64_t* out, void* stream);

/* Number of groups currently in the table (waits for the table's work). */
int qk_gb_group_count(qk_groupby h, int64_t* out, void* stream);

/* Kernel timing: when enabled, HIP events bracket every row-aggregation kernel on its
 * stream; *ms is the summed duration, *launches the number of launches measured. */
int qk_gb_set_profiling(qk_groupby h, int32_t enable);
int qk_gb_get_profile(qk_groupby h, double* ms, int64_t* launches);

/* Describe the table's layout: bytes per global slot, number of slots, LDS slots per
 * workgroup used by the last launch. */
int qk_gb_describe(qk_groupby h, int32_t* slot_bytes, int64_t* slots, int32_t* lds_slots);

/* Launch shape of the row kernel (no reference counterpart; the reference always launches
 * ceil(n/1024) x maxThreadsPerBlock, Context.cu:109-127).  block: threads per workgroup
 * (multiple of 64, <= 1024); lds_bytes: LDS table bytes per workgroup (0 = no LDS stage);
 * wgs_per_cu: resident workgroups per CU (0 = as many as fit).  A negative argument leaves
 * that setting unchanged. */
int qk_gb_set_tuning(qk_groupby h, int32_t block, int32_t lds_bytes, int32_t wgs_per_cu);

/* Row-kernel path for single-word keys (no reference counterpart): path -1 = automatic
 * (two-phase partitioned aggregation for blocks of at least min_rows rows, the one-pass
 * kernel below that), 0 = always one-pass, 1 = always two-phase.  min_rows < 0 leaves the
 * threshold unchanged.  Results do not depend on the path. */
int qk_gb_set_path(qk_groupby h, int32_t path, int64_t min_rows);

/* Deferred merging on the two-phase path (no reference counterpart).  A table created with
 * QK_GB_ASYNC may keep the partitioned rows of several blocks and merge them into the table
 * once, when the table is next read (qk_gb_get_results, qk_gb_group_count,
 * qk_gb_check_error) or when the buffers are full.  rows: how many rows may be kept
 * (-1 = default: four blocks, at most 2^28 rows; 0 = merge after every block).  The
 * partition buffers take about 24 bytes per kept row. */
int qk_gb_set_defer_rows(qk_groupby h, int64_t rows);

/* Internal state for tuning (no reference counterpart); waits for the table's work.
 * out[0] path of the last block (0 one-pass, 1 two-phase), out[1] heavy keys, out[2] heavy
 * count threshold, out[3] rows written to partition regions by the last block, out[4]
 * phase-1 workgroups, out[5] partitions, out[6] rows per region, out[7] phase-1 LDS slots,
 * out[8] phase-2 LDS slots, out[9] heavy key limit, out[10] sub-tables of the table, out[11]
 * growths of a QK_GB_AUTO_GROW table, out[12] phase-1 kernel of the last two-phase block (1
 * streaming form, 0 general form).  n = entries to fill (at most 13). */
int qk_gb_stats(qk_groupby h, int64_t* out, int32_t n, void* stream);

/* ======================================================================================
 * Batched queries (no reference counterpart: the reference runs one query at a time, each
 * block through filter -> reconstruct -> GroupBy, GpuSqlDispatcher.cpp:1107-1179)
 *
 * Many independent `SELECT key, SUM(v), MIN(v), MAX(v), COUNT(v) FROM t WHERE lo <= ts < hi
 * GROUP BY key` queries over one device-resident table, run together by one persistent
 * kernel that takes (tile, query) work units from device queues, one queue per die of the
 * device (the queries of one tile run on one die and share its L2).  Each query reads a
 * contiguous range of the table's blocks and groups by a computed key
 *     key = (a * b_range + b) mod key_mod        (Euclidean remainder, 0 <= key < key_mod)
 * of two int32 columns, so its groups are the dense range [0, key_mod).  A query with at
 * most `lds_groups` groups aggregates each work unit in LDS and adds the unit's groups to
 * its result with atomics.  A larger query takes two passes: its passing rows are first
 * written to a scratch pool grouped by key chunk, then each chunk is aggregated in LDS by one
 * workgroup (rows that find no room in the pool are aggregated with device atomics).
 * ====================================================================================== */

typedef struct qk_batch_columns {
    const int32_t* a;       /* key part: key = (a * b_range + b) mod key_mod */
    const int32_t* b;
    const int64_t* ts;      /* filtered column */
    const int64_t* value;   /* aggregated column (no NULLs) */
    int64_t rows;           /* table rows */
    int32_t block_rows;     /* rows per block */
    int32_t b_range;        /* multiplier of a in the key */
} qk_batch_columns;

typedef struct qk_batch_query {
    int32_t first_block;    /* blocks [first_block, first_block + nblocks) */
    int32_t nblocks;
    int64_t ts_lo, ts_hi;   /* WHERE ts >= ts_lo AND ts < ts_hi */
    int32_t key_mod;        /* groups: 1 .. 2^24 */
    int32_t reserved;       /* 0 */
} qk_batch_query;

typedef struct qk_batch_tuning {
    int32_t tile_rows;      /* rows per work unit (multiple of 4096; its pass bits, tile_rows / 8
                               bytes, and the chunk histogram must fit the LDS) */
    int32_t lds_groups;     /* largest key_mod aggregated in LDS in one pass */
    int32_t wgs_per_cu;     /* persistent workgroups per CU (1 or 2) */
    int32_t use_graph;      /* replay the batch as a captured hipGraph */
    int64_t scratch_rows;   /* most rows of the pool of the two-pass path, 12 B per row (0 = 2^32);
                               it is allocated as large as a batch's estimated rows need */
} qk_batch_tuning;

typedef struct qk_batch_s* qk_batch;

/* A batch engine over one table (columns stay owned by the caller).  Creating it takes sorted
 * samples of ts from every block (16384 per block), from which the passing rows of a query are
 * estimated on the host; qk_batch_resample takes them again after the caller changes ts. */
int qk_batch_create(qk_batch* out, const qk_batch_columns* cols);
int qk_batch_destroy(qk_batch b);
int qk_batch_resample(qk_batch b);
/* Estimated passing rows of each query (host only, no device work). */
int qk_batch_estimate_rows(qk_batch b, int32_t nq, const qk_batch_query* q, int64_t* out);
int qk_batch_set_tuning(qk_batch b, const qk_batch_tuning* t);
int qk_batch_get_tuning(qk_batch b, qk_batch_tuning* t);
/* Time candidate tunings on the given queries (they are run several times) and keep the
 * fastest; *out (may be NULL) receives it. */
int qk_batch_autotune(qk_batch b, int32_t nq, const qk_batch_query* q, qk_batch_tuning* out,
                      void* stream);
/* Run nq queries.  Results stay in the engine until the next run: for query i, groups
 * g in [0, key_mod_i) are at offset off[i] + g of the four device arrays (a group exists
 * when its count is non-zero; MIN/MAX of an empty group are undefined).  Returns after
 * the work is queued; qk_stream_synchronize(stream) waits for it. */
int qk_batch_run(qk_batch b, int32_t nq, const qk_batch_query* q, void* stream);
/* As qk_batch_run, with the results written to the caller's device arrays (each at least
 * sum(key_mod) entries, query i at the i-th prefix sum of key_mod) instead of the engine's;
 * qk_batch_results then reports those arrays. */
int qk_batch_run_into(qk_batch b, int32_t nq, const qk_batch_query* q, int64_t* count, int64_t* sum,
                      int64_t* min, int64_t* max, void* stream);
/* Device arrays of the last run and the host array of offsets (nq + 1 entries). */
int qk_batch_results(qk_batch b, const int64_t** count, const int64_t** sum, const int64_t** min,
                     const int64_t** max, const int64_t** offsets_host);
/* Completion time of each query of the last run, in ms from the start of the run (waits
 * for the run). */
int qk_batch_completion_ms(qk_batch b, double* out, int32_t nq);

/* key[i] = (a[i] * b_range + b[i]) mod key_mod as int64: the key column a one-query-at-a-
 * time caller materialises before GROUP BY. */
int qk_batch_key_column(const int32_t* a, const int32_t* b, int32_t b_range, int32_t key_mod,
                        int32_t n, int64_t* out, void* stream);

/* ======================================================================================
 * Serving (no reference counterpart: the reference gives each client connection a worker
 * thread that runs its queries one at a time, ClientPoolWorker.cpp:42-140)
 *
 * Continuous batching of requests of the batched form above.  Requests are queued as they are
 * submitted; a dispatcher thread runs the oldest ones as a batch on a free engine (one stream
 * each) while the projected bytes of all admitted, unreleased requests stay under memory_cap,
 * and starts the next batch while the previous one is still on the device.  Projected bytes
 * of a request: key_mod x 32 B of dense result table, plus (estimated rows x 1.25 + 4096) x
 * 12 B of two-pass pool when key_mod > tuning.lds_groups.  A request is done when its batch
 * has finished on the device; its groups stay in device memory until it is released, and the
 * engine of a batch takes new work once all the batch's requests are released.
 * ====================================================================================== */

typedef struct qk_serve_config {
    int32_t engines;          /* batch engines: batches running or holding results (0 = 2) */
    int32_t max_batch;        /* most requests per batch (0 = 4096) */
    int64_t memory_cap;       /* projected bytes admitted at once (0 = 200 GB); the engines' row
                                 pools share the device memory left over it (scratch_rows is
                                 lowered to fit, QK_OUT_OF_MEMORY when under 2^20 rows each) */
    double wait_ms;           /* with a batch running, the next one starts when its oldest
                                 request has waited this long (0 = half the last batch's time) */
    qk_batch_tuning tuning;   /* of every engine (tile_rows 0 = defaults, no graph replay) */
} qk_serve_config;

typedef struct qk_serve_stats {
    int64_t batches, served, failed, pending;
    int64_t max_batch;                          /* largest batch run */
    int64_t admitted_bytes, peak_admitted_bytes;
    double busy_ms;                             /* sum of batch times */
} qk_serve_stats;

typedef struct qk_serve_s* qk_serve;

int qk_serve_create(qk_serve* out, const qk_batch_columns* cols, const qk_serve_config* cfg);
/* Waits for the batches on the device; requests not yet started fail. */
int qk_serve_destroy(qk_serve s);
/* Thread-safe.  QK_OUT_OF_MEMORY when the request alone projects more than memory_cap. */
int qk_serve_submit(qk_serve s, const qk_batch_query* q, int64_t* ticket);
/* Waits until the request is done (timeout_ms < 0: no limit); returns its error if it failed. */
int qk_serve_wait(qk_serve s, int64_t ticket, double timeout_ms);
/* Device pointers to groups [0, key_mod) of a done request (as qk_batch_results), valid
 * until qk_serve_release. */
int qk_serve_result(qk_serve s, int64_t ticket, const int64_t** count, const int64_t** sum,
                    const int64_t** min, const int64_t** max);
/* Submit-to-done time of a done request, in ms. */
int qk_serve_latency_ms(qk_serve s, int64_t ticket, double* out);
/* Drops a done or pending request (a running one cannot be released). */
int qk_serve_release(qk_serve s, int64_t ticket);
/* Open-loop replay: submits request i at arrival_ms[i] (ms from the call, non-decreasing),
 * waits for all and stores each one's latency from its scheduled arrival to done. */
int qk_serve_replay(qk_serve s, int32_t n, const qk_batch_query* q, const double* arrival_ms,
                    double* latency_ms);
int qk_serve_get_stats(qk_serve s, qk_serve_stats* out);

/* The plain path under the same open-loop arrivals, for comparison with qk_serve_replay: one
 * worker runs one request at a time to completion, in arrival order, as the reference's client
 * worker runs its queries (TCPClientHandler.cpp:197-245, GpuSqlDispatcher.cpp:1107-1179).  Each
 * request goes through the drop-in per-block path: the block's key column
 * (qk_batch_key_column), the fused WHERE + GROUP BY of a GPUGroupBy table
 * (qk_gb_process_block_where), then GetResults into device memory.  latency_ms[i] is the time
 * from arrival_ms[i] (ms after the call starts) until request i's groups are in device memory. */
int qk_serve_replay_plain(const qk_batch_columns* cols, int32_t n, const qk_batch_query* q,
                          const double* arrival_ms, double* latency_ms, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QIKKDB_AMD_H */
