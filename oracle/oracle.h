/*
 * oracle.h - CPU restatement of qikkDB's WHERE -> reconstruct -> GROUP BY semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative to the qikkDB
 * source tree).  Parity pins (golden vectors transcribed from the reference's own tests,
 * closed forms, invariants) live in tests/test_oracle.py.
 */
#ifndef QIKK_ORACLE_H
#define QIKK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same numbering as the product ABI (DataType.h:18-25, GPUFilterConditions.cuh,
 * AggregationFunctions.cuh). */
enum { OR_INT = 8, OR_LONG = 9, OR_FLOAT = 10, OR_DOUBLE = 11, OR_INT8 = 15 };
enum { OR_GREATER = 0, OR_GREATER_EQUAL, OR_LESS, OR_LESS_EQUAL, OR_EQUAL, OR_NOT_EQUAL,
       OR_LOGICAL_AND, OR_LOGICAL_OR };
enum { OR_COLUMN = 0, OR_CONST = 1 };
enum { OR_AGG_MIN = 0, OR_AGG_MAX, OR_AGG_SUM, OR_AGG_AVG, OR_AGG_COUNT, OR_AGG_NONE,
       OR_AGG_COUNT_ALL };

/* One comparison of a conjunctive WHERE; all pointers are host pointers. */
typedef struct or_pred {
    int32_t op;
    int32_t type;
    const void* column;
    const uint64_t* null_mask;
    int32_t rhs_kind;
    int32_t rhs_type;
    const void* rhs_column;
    const uint64_t* rhs_null_mask;
    int64_t rhs_int;
    double rhs_real;
} or_pred;

/* kernel_filter (GPUFilter.cuh:22-44) + GPUBinary filter launcher (GPUBinary.cuh:565-578).
 * Returns 0, or -1 for an unsupported type/op. */
int or_filter(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
              int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask, int8_t* out);
/* kernel_operator_not (GPULogic.cuh:17-33). */
void or_logical_not(const int8_t* in, int32_t n, const uint64_t* null_mask, int8_t* out);
/* 1 when row `row` passes every predicate (AND), 0 otherwise (GPUFilter.cuh:28-43). */
int or_eval_where(int32_t npreds, const or_pred* preds, int64_t row);

/* GPUReconstruct::PrefixSum / PrefixSumExclusive (GPUReconstruct.cuh:572-601). */
void or_prefix_sum(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive);
/* GPUReconstruct::Sum on a mask (GPUReconstruct.cuh:604-624). */
int64_t or_mask_sum(const int8_t* mask, int32_t n);
/* reconstructColKeep data + null mask (GPUReconstruct.cuh:29-45, :271-355,
 * GPUReconstruct.cu:8-28).  out_null may be NULL; it must hold ceil(n/64) words.
 * Returns the kept count. */
int32_t or_compact(int32_t elem_size, const void* in, const int8_t* mask, int32_t n,
                   const uint64_t* in_null, void* out, uint64_t* out_null);
/* GenerateIndexes (GPUReconstruct.cuh:55-71, :486-566). */
int32_t or_generate_indexes(const int8_t* mask, int32_t n, int32_t* out);
/* CompressNullMask (GPUReconstruct.cu:31-43, :756-764). */
void or_compress_null_mask(const uint8_t* in, int32_t n, uint64_t* out);

/* ---- GROUP BY ------------------------------------------------------------------------- */
typedef struct or_gb or_gb;

/* A group table keeping naggs aggregates (each its own single-aggregate query in the
 * reference, S11).  out_types as in qk_gb_create (NULL = reference typing). */
or_gb* or_gb_create(int32_t nkeys, const int32_t* key_types, int32_t value_type, int32_t naggs,
                    const int32_t* aggs, const int32_t* out_types);
void or_gb_destroy(or_gb* g);
/* Aggregate rows [begin, end).  Row i takes part when mask == NULL or mask[i] != 0, and it
 * passes the npreds predicates.  group_by_kernel (GPUGroupBy.cuh:68-196),
 * kernel_group_by_multi_key (GPUGroupByMultiKey.cuh:156-259). */
void or_gb_process(or_gb* g, const void* const* keys, const uint64_t* const* key_nulls,
                   const void* values, const uint64_t* value_nulls, const int8_t* mask,
                   int32_t npreds, const or_pred* preds, int64_t begin, int64_t end);
/* Fold src's groups into dst (the multi-GPU merge of GPUGroupBy.cuh:594-790). */
int or_gb_merge(or_gb* dst, const or_gb* src);
int64_t or_gb_count(const or_gb* g);
/* Results in unspecified order (S10).  out_keys[k] is a host array of the key type;
 * out_key_nulls[k] one byte per group.  For aggregate a: out_int[a] (integer results) or
 * out_real[a] (floating results: SUM/AVG of FLOAT/DOUBLE are given in FP64) and
 * out_null[a] one byte per group.  Any pointer may be NULL to skip it. */
void or_gb_results(const or_gb* g, void* const* out_keys, uint8_t* const* out_key_nulls,
                   int64_t* const* out_int, double* const* out_real, uint8_t* const* out_null);
/* 1 when aggregate a's result is floating (out_real), 0 when integer (out_int). */
int or_gb_result_is_real(const or_gb* g, int32_t a);

/* The same aggregation run by nthreads threads over row slices, each into its own table,
 * then merged (CPU baseline timing).  Returns the merged table. */
or_gb* or_gb_query_mt(int32_t nthreads, int32_t nkeys, const int32_t* key_types,
                      int32_t value_type, int32_t naggs, const int32_t* aggs,
                      const int32_t* out_types, const void* const* keys,
                      const uint64_t* const* key_nulls, const void* values,
                      const uint64_t* value_nulls, int32_t npreds, const or_pred* preds,
                      int64_t n);

#ifdef __cplusplus
}
#endif

#endif
