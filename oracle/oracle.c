/*
 * oracle.c - CPU restatement of qikkDB's WHERE -> reconstruct -> GROUP BY semantics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline.  The
 * product never links or loads it.  Plain C99, built by oracle/Makefile.
 *
 * Rules restated here (reference paths relative to the qikkDB tree):
 *   S1  a NULL operand makes a comparison false; AND/OR act on 0/1 masks
 *       (QueryEngine/GPUCore/GPUFilter.cuh:28-43, GPUFilterConditions.cuh:9-110).
 *   S2  comparisons use the C/C++ usual arithmetic conversions of the two operand types
 *       (GPUFilterConditions.cuh:11-14); here the C compiler applies them itself.
 *   S3  only rows whose mask/predicates hold reach GROUP BY, in row order
 *       (GPUReconstruct.cuh:29-45, GpuSqlDispatcherAggregationFunctions.h:422-442).
 *   S4  a NULL key forms one group (GPUGroupBy.cuh:162-166); S9 composite keys compare
 *       column-wise with NULL equal to NULL (GPUGroupByMultiKey.cu:140-208).
 *   S5  a NULL value is neither aggregated nor counted by COUNT(col) (GPUGroupBy.cuh:177-193).
 *   S6  a group with no non-NULL value has NULL SUM/MIN/MAX/AVG and COUNT 0
 *       (GPUGroupBy.cuh:329, :561-585).
 *   S7  output types: SUM/MIN/MAX/AVG give the value type (int32 SUM wraps modulo 2^32,
 *       int32 AVG = wrapped int32 sum / int64 count), COUNT gives int64
 *       (GPUGroupBy.cuh:522-559, GPUBinary.cuh:126-143, DispatcherMacros.h:202-250).
 *   S8  MIN starts at max(), MAX at lowest(), SUM/AVG/COUNT at 0 (AggregationFunctions.cuh).
 * Floating SUM/AVG are accumulated in FP64 here: the reference's FP32 atomics have no fixed
 * order (AggregationFunctions.cuh:219-223), so the FP64 value is the reference point.
 * NaN and signed zero, where the reference's result depends on row order, follow a chosen
 * rule instead (N1): a FLOAT/DOUBLE key -0.0 groups with +0.0 and every NaN key falls in one
 * group; float MIN/MAX follow IEEE 754 totalOrder, so a positive NaN is the MAX of its group
 * and is skipped by MIN unless the group holds nothing else (see key_bits and total_lt).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NULLBIT(m, i) ((int)(((m)[(uint64_t)(i) >> 6] >> ((uint64_t)(i)&63u)) & 1u))

/* The FilterConditions functors (GPUFilterConditions.cuh:9-110): `a OP b` on the operand
 * types as given, so the C compiler applies the usual arithmetic conversions (S2). */
#define OR_CMP(op, a, b)                                                                   \
    ((op) == OR_GREATER         ? ((a) > (b))                                              \
     : (op) == OR_GREATER_EQUAL ? ((a) >= (b))                                             \
     : (op) == OR_LESS          ? ((a) < (b))                                              \
     : (op) == OR_LESS_EQUAL    ? ((a) <= (b))                                             \
     : (op) == OR_EQUAL         ? ((a) == (b))                                             \
     : (op) == OR_NOT_EQUAL     ? ((a) != (b))                                             \
     : (op) == OR_LOGICAL_AND   ? ((a) && (b))                                             \
                                : ((a) || (b)))

#define TKEY(lt, rt) ((lt)*32 + (rt))
#define CMP_CASE(LTC, LT, RTC, RT)                                                         \
    case TKEY(LTC, RTC): {                                                                 \
        LT a = ((const LT*)lp)[li];                                                        \
        RT b = ((const RT*)rp)[ri];                                                        \
        return OR_CMP(op, a, b) ? 1 : 0;                                                   \
    }
#define CMP_ROW(LTC, LT)                                                                   \
    CMP_CASE(LTC, LT, OR_INT8, int8_t)                                                     \
    CMP_CASE(LTC, LT, OR_INT, int32_t)                                                     \
    CMP_CASE(LTC, LT, OR_LONG, int64_t)                                                    \
    CMP_CASE(LTC, LT, OR_FLOAT, float)                                                     \
    CMP_CASE(LTC, LT, OR_DOUBLE, double)

/* Compare element li of lp (type lt) with element ri of rp (type rt); -1 = bad types. */
static int cmp_values(int op, int lt, const void* lp, int64_t li, int rt, const void* rp,
                      int64_t ri) {
    switch (TKEY(lt, rt)) {
        CMP_ROW(OR_INT8, int8_t)
        CMP_ROW(OR_INT, int32_t)
        CMP_ROW(OR_LONG, int64_t)
        CMP_ROW(OR_FLOAT, float)
        CMP_ROW(OR_DOUBLE, double)
    default:
        return -1;
    }
}

static int type_ok(int t) {
    return t == OR_INT8 || t == OR_INT || t == OR_LONG || t == OR_FLOAT || t == OR_DOUBLE;
}

/* A scalar operand keeps its literal type (S2): stored in `buf` as that type. */
static const void* scalar_from_pred(int type, int64_t iv, double rv, void* buf) {
    switch (type) {
    case OR_INT8: *(int8_t*)buf = (int8_t)iv; break;
    case OR_INT: *(int32_t*)buf = (int32_t)iv; break;
    case OR_LONG: *(int64_t*)buf = iv; break;
    case OR_FLOAT: *(float*)buf = (float)rv; break;
    default: *(double*)buf = rv; break;
    }
    return buf;
}

int or_filter(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
              int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask,
              int8_t* out) {
    if (!type_ok(ltype) || !type_ok(rtype) || op < 0 || op > OR_LOGICAL_OR) return -1;
    for (int32_t i = 0; i < n; i++) {
        int r = cmp_values(op, ltype, lhs, lkind == OR_COLUMN ? i : 0, rtype, rhs,
                           rkind == OR_COLUMN ? i : 0);
        if (null_mask && NULLBIT(null_mask, i)) r = 0; /* GPUFilter.cuh:32-35 */
        out[i] = (int8_t)r;
    }
    return 0;
}

void or_logical_not(const int8_t* in, int32_t n, const uint64_t* null_mask, int8_t* out) {
    for (int32_t i = 0; i < n; i++)
        out[i] = (int8_t)(!in[i] && !(null_mask && NULLBIT(null_mask, i)));
}

int or_eval_where(int32_t npreds, const or_pred* preds, int64_t row) {
    for (int32_t p = 0; p < npreds; p++) {
        const or_pred* q = &preds[p];
        if (q->null_mask && NULLBIT(q->null_mask, row)) return 0;
        int r;
        if (q->rhs_kind == OR_CONST) {
            double buf;
            const void* c = scalar_from_pred(q->rhs_type, q->rhs_int, q->rhs_real, &buf);
            r = cmp_values(q->op, q->type, q->column, row, q->rhs_type, c, 0);
        } else {
            if (q->rhs_null_mask && NULLBIT(q->rhs_null_mask, row)) return 0;
            r = cmp_values(q->op, q->type, q->column, row, q->rhs_type, q->rhs_column, row);
        }
        if (r <= 0) return 0;
    }
    return 1;
}

/* ---- reconstruct ------------------------------------------------------------------- */

void or_prefix_sum(const int8_t* mask, int32_t n, int32_t* out, int32_t inclusive) {
    int32_t acc = 0;
    for (int32_t i = 0; i < n; i++) {
        if (inclusive) {
            acc += mask[i];
            out[i] = acc;
        } else {
            out[i] = acc;
            acc += mask[i];
        }
    }
}

int64_t or_mask_sum(const int8_t* mask, int32_t n) {
    int64_t s = 0;
    for (int32_t i = 0; i < n; i++) s += mask[i];
    return s;
}

int32_t or_compact(int32_t elem_size, const void* in, const int8_t* mask, int32_t n,
                   const uint64_t* in_null, void* out, uint64_t* out_null) {
    int32_t k = 0;
    if (out_null) memset(out_null, 0, (size_t)((n + 63) / 64) * 8);
    for (int32_t i = 0; i < n; i++) {
        if (!mask[i]) continue;
        memcpy((char*)out + (size_t)k * elem_size, (const char*)in + (size_t)i * elem_size,
               (size_t)elem_size);
        if (out_null && in_null && NULLBIT(in_null, i))
            out_null[k >> 6] |= (uint64_t)1 << (k & 63);
        k++;
    }
    return k;
}

int32_t or_generate_indexes(const int8_t* mask, int32_t n, int32_t* out) {
    int32_t k = 0;
    for (int32_t i = 0; i < n; i++)
        if (mask[i]) out[k++] = i;
    return k;
}

void or_compress_null_mask(const uint8_t* in, int32_t n, uint64_t* out) {
    memset(out, 0, (size_t)((n + 63) / 64) * 8);
    for (int32_t i = 0; i < n; i++)
        out[i >> 6] |= (uint64_t)(in[i] & 1u) << (i & 63);
}

/* ---- GROUP BY ---------------------------------------------------------------------- */

#define OR_MAX_KEYS 8
#define OR_MAX_AGGS 16

struct or_gb {
    int32_t nkeys, value_type, naggs;
    int32_t key_types[OR_MAX_KEYS];
    int32_t aggs[OR_MAX_AGGS], out_types[OR_MAX_AGGS];
    int64_t cap, count;
    uint64_t* kbits; /* cap * nkeys key bit patterns (NULL column -> 0) */
    uint8_t* knull;  /* cap * nkeys */
    uint8_t* used;
    int64_t *cnt_nn, *cnt_all;
    uint64_t* sum_i; /* wrapping integer sum */
    double* sum_f;
    int64_t *min_i, *max_i;
    double *min_f, *max_f;
};

static int is_real(int t) { return t == OR_FLOAT || t == OR_DOUBLE; }

/* Key bit pattern used for equality: integers sign-extended, floats by value with -0.0
 * folded into +0.0 so that `==` semantics hold for zeros, and every NaN folded into one
 * quiet NaN so that NaN keys form one group (a deliberate, deterministic departure from
 * GPUGroupBy.cuh:105, whose `!=` key test gives each NaN row a group of its own). */
static uint64_t key_bits(int type, const void* col, int64_t i) {
    switch (type) {
    case OR_INT8: return (uint64_t)(int64_t)((const int8_t*)col)[i];
    case OR_INT: return (uint64_t)(int64_t)((const int32_t*)col)[i];
    case OR_LONG: return (uint64_t)((const int64_t*)col)[i];
    case OR_FLOAT: {
        float f = ((const float*)col)[i];
        uint32_t b;
        if (f == 0.0f) f = 0.0f;
        if (f != f) return 0x7fc00000u;
        memcpy(&b, &f, 4);
        return b;
    }
    default: {
        double d = ((const double*)col)[i];
        uint64_t b;
        if (d == 0.0) d = 0.0;
        if (d != d) return 0x7ff8000000000000ull;
        memcpy(&b, &d, 8);
        return b;
    }
    }
}

/* Strict "a before b" in IEEE 754-2019 totalOrder (clause 5.10): -NaN < -inf < ... < -0.0 <
 * +0.0 < ... < +inf < +NaN, NaNs of one sign ordered by payload.  Float MIN/MAX use this order:
 * the reference's CAS loops (AggregationFunctions.cuh:44-80 and their max twins) restart from a
 * NaN, so their result depends on row order and no parallel order can reproduce it. */
static int64_t total_rank(double d) {
    int64_t b;
    memcpy(&b, &d, 8);
    return b < 0 ? (int64_t)((uint64_t)b ^ 0x7fffffffffffffffull) : b;
}
static int total_lt(double a, double b) { return total_rank(a) < total_rank(b); }
static double bits_double(uint64_t u) {
    double d;
    memcpy(&d, &u, 8);
    return d;
}
/* The first and last values in totalOrder (a NaN of each sign with the largest payload): the
 * MAX and MIN states of a group that has seen no value yet. */
#define total_first bits_double(0xffffffffffffffffull)
#define total_last bits_double(0x7fffffffffffffffull)

/* FNV-1a over the key words (a different hash from the GPU table on purpose). */
static uint64_t hash_key(int nk, const uint64_t* bits, const uint8_t* nulls) {
    uint64_t h = 1469598103934665603ull;
    for (int k = 0; k < nk; k++) {
        uint64_t w = bits[k] ^ ((uint64_t)nulls[k] << 63) ^ ((uint64_t)nulls[k] << 17);
        for (int b = 0; b < 8; b++) {
            h ^= (w >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    }
    h ^= h >> 29;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 32;
    return h;
}

static int alloc_arrays(or_gb* g, int64_t cap) {
    int nk = g->nkeys;
    g->cap = cap;
    g->kbits = (uint64_t*)calloc((size_t)(cap * nk), 8);
    g->knull = (uint8_t*)calloc((size_t)(cap * nk), 1);
    g->used = (uint8_t*)calloc((size_t)cap, 1);
    g->cnt_nn = (int64_t*)calloc((size_t)cap, 8);
    g->cnt_all = (int64_t*)calloc((size_t)cap, 8);
    g->sum_i = (uint64_t*)calloc((size_t)cap, 8);
    g->sum_f = (double*)calloc((size_t)cap, 8);
    g->min_i = (int64_t*)malloc((size_t)cap * 8);
    g->max_i = (int64_t*)malloc((size_t)cap * 8);
    g->min_f = (double*)malloc((size_t)cap * 8);
    g->max_f = (double*)malloc((size_t)cap * 8);
    if (!g->kbits || !g->knull || !g->used || !g->cnt_nn || !g->cnt_all || !g->sum_i ||
        !g->sum_f || !g->min_i || !g->max_i || !g->min_f || !g->max_f)
        return -1;
    for (int64_t s = 0; s < cap; s++) {
        g->min_i[s] = INT64_MAX;
        g->max_i[s] = INT64_MIN;
        g->min_f[s] = total_last;
        g->max_f[s] = total_first;
    }
    return 0;
}

static void free_arrays(or_gb* g) {
    free(g->kbits); free(g->knull); free(g->used); free(g->cnt_nn); free(g->cnt_all);
    free(g->sum_i); free(g->sum_f); free(g->min_i); free(g->max_i); free(g->min_f);
    free(g->max_f);
}

or_gb* or_gb_create(int32_t nkeys, const int32_t* key_types, int32_t value_type, int32_t naggs,
                    const int32_t* aggs, const int32_t* out_types) {
    if (nkeys < 1 || nkeys > OR_MAX_KEYS || naggs < 1 || naggs > OR_MAX_AGGS) return NULL;
    or_gb* g = (or_gb*)calloc(1, sizeof(or_gb));
    if (!g) return NULL;
    g->nkeys = nkeys;
    g->value_type = value_type;
    g->naggs = naggs;
    for (int k = 0; k < nkeys; k++) g->key_types[k] = key_types[k];
    for (int a = 0; a < naggs; a++) {
        g->aggs[a] = aggs[a];
        if (out_types && out_types[a] > 0)
            g->out_types[a] = out_types[a];
        else
            g->out_types[a] = (aggs[a] == OR_AGG_COUNT || aggs[a] == OR_AGG_COUNT_ALL)
                                  ? OR_LONG : value_type;
    }
    if (alloc_arrays(g, 64) != 0) {
        or_gb_destroy(g);
        return NULL;
    }
    return g;
}

void or_gb_destroy(or_gb* g) {
    if (!g) return;
    free_arrays(g);
    free(g);
}

static int64_t find_slot(const or_gb* g, const uint64_t* bits, const uint8_t* nulls,
                         int* found) {
    int nk = g->nkeys;
    uint64_t mask = (uint64_t)g->cap - 1;
    uint64_t s = hash_key(nk, bits, nulls) & mask;
    for (;;) {
        if (!g->used[s]) {
            *found = 0;
            return (int64_t)s;
        }
        int eq = 1;
        for (int k = 0; k < nk && eq; k++) {
            uint8_t n1 = g->knull[s * nk + k];
            if (n1 != nulls[k]) eq = 0;
            else if (!n1 && g->kbits[s * nk + k] != bits[k]) eq = 0;
        }
        if (eq) {
            *found = 1;
            return (int64_t)s;
        }
        s = (s + 1) & mask;
    }
}

static int grow(or_gb* g) {
    or_gb old = *g;
    if (alloc_arrays(g, old.cap * 2) != 0) return -1;
    int nk = g->nkeys;
    for (int64_t s = 0; s < old.cap; s++) {
        if (!old.used[s]) continue;
        int found;
        int64_t d = find_slot(g, &old.kbits[s * nk], &old.knull[s * nk], &found);
        g->used[d] = 1;
        memcpy(&g->kbits[d * nk], &old.kbits[s * nk], (size_t)nk * 8);
        memcpy(&g->knull[d * nk], &old.knull[s * nk], (size_t)nk);
        g->cnt_nn[d] = old.cnt_nn[s]; g->cnt_all[d] = old.cnt_all[s];
        g->sum_i[d] = old.sum_i[s]; g->sum_f[d] = old.sum_f[s];
        g->min_i[d] = old.min_i[s]; g->max_i[d] = old.max_i[s];
        g->min_f[d] = old.min_f[s]; g->max_f[d] = old.max_f[s];
    }
    free_arrays(&old);
    return 0;
}

static int64_t find_or_insert(or_gb* g, const uint64_t* bits, const uint8_t* nulls) {
    if ((g->count + 1) * 2 > g->cap)
        if (grow(g) != 0) return -1;
    int found;
    int64_t s = find_slot(g, bits, nulls, &found);
    if (!found) {
        int nk = g->nkeys;
        g->used[s] = 1;
        memcpy(&g->kbits[s * nk], bits, (size_t)nk * 8);
        memcpy(&g->knull[s * nk], nulls, (size_t)nk);
        g->count++;
    }
    return s;
}

void or_gb_process(or_gb* g, const void* const* keys, const uint64_t* const* key_nulls,
                   const void* values, const uint64_t* value_nulls, const int8_t* mask,
                   int32_t npreds, const or_pred* preds, int64_t begin, int64_t end) {
    int nk = g->nkeys;
    uint64_t bits[OR_MAX_KEYS];
    uint8_t nulls[OR_MAX_KEYS];
    int vt = g->value_type;
    for (int64_t i = begin; i < end; i++) {
        if (mask && !mask[i]) continue;
        if (npreds > 0 && !or_eval_where(npreds, preds, i)) continue;
        for (int k = 0; k < nk; k++) {
            int isnull = key_nulls && key_nulls[k] && NULLBIT(key_nulls[k], i);
            nulls[k] = (uint8_t)isnull;
            bits[k] = isnull ? 0 : key_bits(g->key_types[k], keys[k], i);
        }
        int64_t s = find_or_insert(g, bits, nulls);
        if (s < 0) return;
        g->cnt_all[s]++;
        if (value_nulls && NULLBIT(value_nulls, i)) continue; /* S5 */
        g->cnt_nn[s]++;
        if (!values) continue;
        switch (vt) {
        case OR_INT:
        case OR_LONG:
        case OR_INT8: {
            int64_t v = vt == OR_INT ? ((const int32_t*)values)[i]
                        : vt == OR_LONG ? ((const int64_t*)values)[i]
                                        : ((const int8_t*)values)[i];
            g->sum_i[s] += (uint64_t)v;
            if (v < g->min_i[s]) g->min_i[s] = v;
            if (v > g->max_i[s]) g->max_i[s] = v;
            break;
        }
        default: {
            double v = vt == OR_FLOAT ? (double)((const float*)values)[i]
                                      : ((const double*)values)[i];
            g->sum_f[s] += v;
            if (total_lt(v, g->min_f[s])) g->min_f[s] = v;
            if (total_lt(g->max_f[s], v)) g->max_f[s] = v;
            break;
        }
        }
    }
}

int or_gb_merge(or_gb* dst, const or_gb* src) {
    int nk = src->nkeys;
    for (int64_t s = 0; s < src->cap; s++) {
        if (!src->used[s]) continue;
        int64_t d = find_or_insert(dst, &src->kbits[s * nk], &src->knull[s * nk]);
        if (d < 0) return -1;
        dst->cnt_nn[d] += src->cnt_nn[s];
        dst->cnt_all[d] += src->cnt_all[s];
        dst->sum_i[d] += src->sum_i[s];
        dst->sum_f[d] += src->sum_f[s];
        if (src->min_i[s] < dst->min_i[d]) dst->min_i[d] = src->min_i[s];
        if (src->max_i[s] > dst->max_i[d]) dst->max_i[d] = src->max_i[s];
        if (total_lt(src->min_f[s], dst->min_f[d])) dst->min_f[d] = src->min_f[s];
        if (total_lt(dst->max_f[d], src->max_f[s])) dst->max_f[d] = src->max_f[s];
    }
    return 0;
}

int64_t or_gb_count(const or_gb* g) { return g->count; }

int or_gb_result_is_real(const or_gb* g, int32_t a) {
    int ag = g->aggs[a];
    if (ag == OR_AGG_COUNT || ag == OR_AGG_COUNT_ALL || ag == OR_AGG_NONE) return 0;
    return is_real(g->value_type);
}

static void put_key(int type, void* out, int64_t r, uint64_t bits) {
    switch (type) {
    case OR_INT8: ((int8_t*)out)[r] = (int8_t)(int64_t)bits; break;
    case OR_INT: ((int32_t*)out)[r] = (int32_t)(int64_t)bits; break;
    case OR_LONG: ((int64_t*)out)[r] = (int64_t)bits; break;
    case OR_FLOAT: {
        uint32_t b = (uint32_t)bits;
        memcpy((float*)out + r, &b, 4);
        break;
    }
    default: memcpy((double*)out + r, &bits, 8); break;
    }
}

/* Integer result of aggregate a for slot s (S6-S8 payloads for NULL groups). */
static int64_t int_result(const or_gb* g, int a, int64_t s) {
    int ag = g->aggs[a], vt = g->value_type, ot = g->out_types[a];
    int64_t c = g->cnt_nn[s];
    switch (ag) {
    case OR_AGG_COUNT: return c;
    case OR_AGG_COUNT_ALL: return g->cnt_all[s];
    case OR_AGG_SUM:
        if (c == 0) return 0;
        if (vt == OR_INT && ot == OR_INT) return (int64_t)(int32_t)(uint32_t)g->sum_i[s];
        if (vt == OR_INT8 && ot == OR_INT8) return (int64_t)(int8_t)(uint8_t)g->sum_i[s];
        return (int64_t)g->sum_i[s];
    case OR_AGG_MIN:
        if (c == 0) return vt == OR_INT ? INT32_MAX : vt == OR_INT8 ? INT8_MAX : INT64_MAX;
        return g->min_i[s];
    case OR_AGG_MAX:
        if (c == 0) return vt == OR_INT ? INT32_MIN : vt == OR_INT8 ? INT8_MIN : INT64_MIN;
        return g->max_i[s];
    case OR_AGG_AVG:
        /* sum / count with the sum in the value type (GPUGroupBy.cuh:536-558,
         * GPUBinary.cuh:126-143); count 0 gives the NULL constant min() (NullConstants.cuh). */
        if (vt == OR_INT) {
            if (c == 0) return INT32_MIN;
            return (int64_t)(int32_t)((int64_t)(int32_t)(uint32_t)g->sum_i[s] / c);
        }
        if (vt == OR_INT8) {
            if (c == 0) return INT8_MIN;
            return (int64_t)(int8_t)((int64_t)(int8_t)(uint8_t)g->sum_i[s] / c);
        }
        if (c == 0) return INT64_MIN;
        return (int64_t)g->sum_i[s] / c;
    default: return 0;
    }
}

static double real_result(const or_gb* g, int a, int64_t s) {
    int ag = g->aggs[a], vt = g->value_type;
    int64_t c = g->cnt_nn[s];
    double big = vt == OR_FLOAT ? 3.4028234663852886e38 : 1.7976931348623157e308;
    switch (ag) {
    case OR_AGG_SUM: return c == 0 ? 0.0 : g->sum_f[s];
    case OR_AGG_MIN: return c == 0 ? big : g->min_f[s];
    case OR_AGG_MAX: return c == 0 ? -big : g->max_f[s];
    case OR_AGG_AVG: return c == 0 ? NAN : g->sum_f[s] / (double)c;
    default: return 0.0;
    }
}

void or_gb_results(const or_gb* g, void* const* out_keys, uint8_t* const* out_key_nulls,
                   int64_t* const* out_int, double* const* out_real, uint8_t* const* out_null) {
    int nk = g->nkeys;
    int64_t r = 0;
    for (int64_t s = 0; s < g->cap; s++) {
        if (!g->used[s]) continue;
        for (int k = 0; k < nk; k++) {
            if (out_keys && out_keys[k])
                put_key(g->key_types[k], out_keys[k], r, g->kbits[s * nk + k]);
            if (out_key_nulls && out_key_nulls[k]) out_key_nulls[k][r] = g->knull[s * nk + k];
        }
        for (int a = 0; a < g->naggs; a++) {
            int ag = g->aggs[a];
            if (ag == OR_AGG_NONE) continue;
            if (or_gb_result_is_real(g, a)) {
                if (out_real && out_real[a]) out_real[a][r] = real_result(g, a, s);
            } else {
                if (out_int && out_int[a]) out_int[a][r] = int_result(g, a, s);
            }
            if (out_null && out_null[a])
                out_null[a][r] = (uint8_t)(ag != OR_AGG_COUNT && ag != OR_AGG_COUNT_ALL &&
                                           g->cnt_nn[s] == 0);
        }
        r++;
    }
}

/* ---- threaded run (CPU baseline) ---------------------------------------------------- */

typedef struct {
    or_gb* g;
    const void* const* keys;
    const uint64_t* const* key_nulls;
    const void* values;
    const uint64_t* value_nulls;
    int32_t npreds;
    const or_pred* preds;
    int64_t begin, end;
} or_job;

static void* run_job(void* p) {
    or_job* j = (or_job*)p;
    or_gb_process(j->g, j->keys, j->key_nulls, j->values, j->value_nulls, NULL, j->npreds,
                  j->preds, j->begin, j->end);
    return NULL;
}

or_gb* or_gb_query_mt(int32_t nthreads, int32_t nkeys, const int32_t* key_types,
                      int32_t value_type, int32_t naggs, const int32_t* aggs,
                      const int32_t* out_types, const void* const* keys,
                      const uint64_t* const* key_nulls, const void* values,
                      const uint64_t* value_nulls, int32_t npreds, const or_pred* preds,
                      int64_t n) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    or_job jobs[256];
    pthread_t tid[256];
    /* slices aligned to 64 rows so that null-mask words are never shared */
    int64_t per = ((n + nthreads - 1) / nthreads + 63) / 64 * 64;
    int started = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t].g = or_gb_create(nkeys, key_types, value_type, naggs, aggs, out_types);
        jobs[t].keys = keys;
        jobs[t].key_nulls = key_nulls;
        jobs[t].values = values;
        jobs[t].value_nulls = value_nulls;
        jobs[t].npreds = npreds;
        jobs[t].preds = preds;
        jobs[t].begin = (int64_t)t * per < n ? (int64_t)t * per : n;
        jobs[t].end = (int64_t)(t + 1) * per < n ? (int64_t)(t + 1) * per : n;
    }
    for (int t = 0; t < nthreads; t++) {
        if (!jobs[t].g) break;
        if (pthread_create(&tid[t], NULL, run_job, &jobs[t]) != 0) break;
        started++;
    }
    for (int t = 0; t < started; t++) pthread_join(tid[t], NULL);
    or_gb* out = jobs[0].g;
    for (int t = 1; t < nthreads; t++) {
        if (out && jobs[t].g) or_gb_merge(out, jobs[t].g);
        or_gb_destroy(jobs[t].g);
    }
    if (started != nthreads) {
        or_gb_destroy(out);
        return NULL;
    }
    return out;
}
