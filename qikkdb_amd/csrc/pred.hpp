// pred.hpp - device-side comparison of two typed operands, the FilterConditions functors
// (QueryEngine/GPUCore/GPUFilterConditions.cuh:9-110) evaluated on the operand types as
// given so that the C++ usual arithmetic conversions apply (S2), and the predicate form
// used inside the fused kernels (a NULL operand fails the row, GPUFilter.cuh:28-43).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace qk {

template <typename A, typename B>
__host__ __device__ __forceinline__ bool cmp_op(int op, A a, B b) {
    switch (op) {
    case QK_GREATER: return a > b;
    case QK_GREATER_EQUAL: return a >= b;
    case QK_LESS: return a < b;
    case QK_LESS_EQUAL: return a <= b;
    case QK_EQUAL: return a == b;
    case QK_NOT_EQUAL: return a != b;
    case QK_LOGICAL_AND: return a && b;
    default: return a || b;
    }
}

template <typename T>
__device__ __forceinline__ T const_as(const qk_predicate& p) {
    if constexpr (std::is_floating_point<T>::value)
        return (T)p.rhs_real;
    else
        return (T)p.rhs_int;
}

template <typename L, typename R>
__device__ __forceinline__ bool pred_typed(const qk_predicate& p, int64_t row) {
    L a = ((const L*)p.column)[row];
    R b = (p.rhs_kind == QK_CONST) ? const_as<R>(p) : ((const R*)p.rhs_column)[row];
    return cmp_op(p.op, a, b);
}

template <typename L>
__device__ __forceinline__ bool pred_rhs(const qk_predicate& p, int64_t row) {
    switch (p.rhs_type) {
    case QK_INT8: return pred_typed<L, int8_t>(p, row);
    case QK_INT: return pred_typed<L, int32_t>(p, row);
    case QK_LONG: return pred_typed<L, int64_t>(p, row);
    case QK_FLOAT: return pred_typed<L, float>(p, row);
    default: return pred_typed<L, double>(p, row);
    }
}

// One predicate on one row.
__device__ __forceinline__ bool eval_pred(const qk_predicate& p, int64_t row) {
    if (p.null_mask && null_bit(p.null_mask, row)) return false;
    if (p.rhs_kind == QK_COLUMN && p.rhs_null_mask && null_bit(p.rhs_null_mask, row)) return false;
    switch (p.type) {
    case QK_INT8: return pred_rhs<int8_t>(p, row);
    case QK_INT: return pred_rhs<int32_t>(p, row);
    case QK_LONG: return pred_rhs<int64_t>(p, row);
    case QK_FLOAT: return pred_rhs<float>(p, row);
    default: return pred_rhs<double>(p, row);
    }
}

// A predicate whose left operand is a value already held in a register (the fused GROUP BY
// kernel compares the value column it has loaded anyway).
template <typename V>
__device__ __forceinline__ bool pred_on_value(const qk_predicate& p, V a, int64_t row) {
    if (p.rhs_kind == QK_COLUMN && p.rhs_null_mask && null_bit(p.rhs_null_mask, row)) return false;
    switch (p.rhs_type) {
    case QK_INT8: return cmp_op(p.op, a, p.rhs_kind == QK_CONST ? const_as<int8_t>(p) : ((const int8_t*)p.rhs_column)[row]);
    case QK_INT: return cmp_op(p.op, a, p.rhs_kind == QK_CONST ? const_as<int32_t>(p) : ((const int32_t*)p.rhs_column)[row]);
    case QK_LONG: return cmp_op(p.op, a, p.rhs_kind == QK_CONST ? const_as<int64_t>(p) : ((const int64_t*)p.rhs_column)[row]);
    case QK_FLOAT: return cmp_op(p.op, a, p.rhs_kind == QK_CONST ? const_as<float>(p) : ((const float*)p.rhs_column)[row]);
    default: return cmp_op(p.op, a, p.rhs_kind == QK_CONST ? const_as<double>(p) : ((const double*)p.rhs_column)[row]);
    }
}

// A conjunction of comparisons of the value column with constants, folded into one interval
// per comparison domain: the type both operands convert to under the usual arithmetic
// conversions (i64 for two integer types, else f32 or f64).  Comparing in that domain is
// what each predicate does, so the fold is exact.  NOT_EQUAL and the logical operators do
// not fold.
//
// Each floating bound is a pair: x passes the lower bound when x > lo_x or x >= lo_i, where
// the unused member is NaN (no x compares true with NaN), so strictness needs no flag.  An
// unused domain holds [-inf, +inf]; that only fails NaN, which fails some folded predicate
// anyway.  Integer bounds are inclusive (a strict bound moves by one; an empty interval has
// lo > hi).
struct ValueRange {
    int32_t doms;  // domains holding a bound: bit 0 i64, bit 1 f32, bit 2 f64
    int64_t lo_i, hi_i;
    float lo_fx, lo_fi, hi_fx, hi_fi;
    double lo_dx, lo_di, hi_dx, hi_di;
};

template <typename T>
__device__ __forceinline__ bool in_bounds(T x, T lo_x, T lo_i, T hi_x, T hi_i) {
    return ((x > lo_x) | (x >= lo_i)) & ((x < hi_x) | (x <= hi_i));
}

template <typename VT>
__device__ __forceinline__ bool in_range(const ValueRange& r, VT v) {
    bool ok = true;
    if constexpr (std::is_integral<VT>::value) ok = ((int64_t)v >= r.lo_i) & ((int64_t)v <= r.hi_i);
    if constexpr (std::is_same<VT, float>::value) {
        ok = in_bounds<float>(v, r.lo_fx, r.lo_fi, r.hi_fx, r.hi_fi);
    } else if constexpr (!std::is_same<VT, double>::value) {
        if (r.doms & 2) ok = ok & in_bounds<float>((float)v, r.lo_fx, r.lo_fi, r.hi_fx, r.hi_fi);
    }
    if (r.doms & 4) ok = ok & in_bounds<double>((double)v, r.lo_dx, r.lo_di, r.hi_dx, r.hi_di);
    return ok;
}

// Predicates travel to kernels by value.
constexpr int kMaxPreds = 8;
struct PredSet {
    int32_t n;
    qk_predicate p[kMaxPreds];
    ValueRange vr;  // the predicates folded, when they all fold (see fold_value_range)
    // fold_column_range: the one integer column every predicate compares
    const void* rcol;
    const uint64_t* rnull;
    int32_t rwide;  // 1: int64, 0: int32
};

// The folded range column at row r, widened to int64.
__device__ __forceinline__ int64_t range_col(const PredSet& ps, int64_t r) {
    return ps.rwide ? __builtin_nontemporal_load((const int64_t*)ps.rcol + r)
                    : (int64_t)__builtin_nontemporal_load((const int32_t*)ps.rcol + r);
}

// Whether a non-NULL value of the range column lies in the folded interval.
__device__ __forceinline__ bool col_in_range(const PredSet& ps, int64_t c) {
    return ps.rwide ? in_range<int64_t>(ps.vr, c) : in_range<int32_t>(ps.vr, (int32_t)c);
}

// Host-side validation of a predicate array; returns QK_SUCCESS or an error.
int check_preds(int32_t npreds, const qk_predicate* preds, PredSet* out);

// Fold the predicates into ps->vr when every one compares the value column (vpred has a bit
// per such predicate) of type vtype with a constant by >, >=, <, <= or ==.  Returns whether
// it did; the fused kernels then test the interval instead of the predicate list.
bool fold_value_range(PredSet* ps, uint32_t vpred, int32_t vtype);

// Fold the predicates into ps->vr / rcol / rnull / rwide when none reads the value column
// (vpred == 0) and every one compares the same int32 or int64 column (same null mask) with a
// constant by >, >=, <, <= or ==.  Returns whether it did; the fused kernels then load that
// column with the rest of a tile and test the interval (a NULL value fails it).
bool fold_column_range(PredSet* ps, uint32_t vpred);

}  // namespace qk
