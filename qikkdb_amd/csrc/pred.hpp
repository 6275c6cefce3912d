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

// Predicates travel to kernels by value.
constexpr int kMaxPreds = 8;
struct PredSet {
    int32_t n;
    qk_predicate p[kMaxPreds];
};

// Host-side validation of a predicate array; returns QK_SUCCESS or an error.
int check_preds(int32_t npreds, const qk_predicate* preds, PredSet* out);

}  // namespace qk
