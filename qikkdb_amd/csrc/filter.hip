// filter.hip - WHERE predicates to masks.
//
// Counterpart of kernel_filter + the GPUBinary filter launcher
// (QueryEngine/GPUCore/GPUFilter.cuh:22-44, GPUBinary.cuh:554-579) and kernel_operator_not
// (GPULogic.cuh:17-33).  The int8 form keeps the reference's mask format (one byte per row);
// the bit form packs 64 rows per uint64 word with one wavefront ballot per 64 rows.
#include <cstring>

#include <cmath>
#include <cstdint>

#include "pred.hpp"

namespace qk {

namespace {

constexpr int kBlock = 256;

int grid_for(int64_t items, int per_block) {
    const DeviceInfo& d = device_info();
    int64_t g = (items + per_block - 1) / per_block;
    int64_t cap = (int64_t)(d.cus > 0 ? d.cus : 64) * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// Each thread handles groups of 4 consecutive rows and stores their 4 mask bytes at once.
template <typename L, typename R>
__global__ void __launch_bounds__(kBlock) k_filter_i8(int op, const L* __restrict__ a, L ac, int acol,
                                                      const R* __restrict__ b, R bc, int bcol,
                                                      const uint64_t* __restrict__ nm,
                                                      int8_t* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t groups = (n + 3) >> 2;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
        const int64_t r0 = g << 2;
        if (r0 + 4 <= n) {
            uint32_t nbits = nm ? (uint32_t)(nm[r0 >> 6] >> (r0 & 63)) : 0u;
            uint32_t packed = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                L x = acol ? a[r0 + j] : ac;
                R y = bcol ? b[r0 + j] : bc;
                uint32_t v = (cmp_op(op, x, y) && !((nbits >> j) & 1u)) ? 1u : 0u;
                packed |= v << (8 * j);
            }
            *reinterpret_cast<uint32_t*>(out + r0) = packed;
        } else {
            for (int64_t r = r0; r < n; r++) {
                L x = acol ? a[r] : ac;
                R y = bcol ? b[r] : bc;
                out[r] = (int8_t)(cmp_op(op, x, y) && !(nm && null_bit(nm, r)));
            }
        }
    }
}

// Scalar form for output pointers that are not 4-byte aligned.
template <typename L, typename R>
__global__ void __launch_bounds__(kBlock) k_filter_i8_unaligned(int op, const L* __restrict__ a, L ac,
                                                                int acol, const R* __restrict__ b,
                                                                R bc, int bcol,
                                                                const uint64_t* __restrict__ nm,
                                                                int8_t* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        L x = acol ? a[r] : ac;
        R y = bcol ? b[r] : bc;
        out[r] = (int8_t)(cmp_op(op, x, y) && !(nm && null_bit(nm, r)));
    }
}

// One row per lane; a ballot packs the wave's 64 results into one word.
template <typename L, typename R>
__global__ void __launch_bounds__(kBlock) k_filter_bits(int op, const L* __restrict__ a, L ac, int acol,
                                                        const R* __restrict__ b, R bc, int bcol,
                                                        const uint64_t* __restrict__ nm,
                                                        uint64_t* __restrict__ out, int64_t n) {
    const int lane = threadIdx.x & 63;
    const int64_t words = (n + 63) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = wave; w < words; w += nwaves) {
        const int64_t r = (w << 6) + lane;
        bool v = false;
        if (r < n) {
            L x = acol ? a[r] : ac;
            R y = bcol ? b[r] : bc;
            v = cmp_op(op, x, y) && !(nm && null_bit(nm, r));
        }
        const uint64_t bits = __ballot(v);
        if (lane == 0) out[w] = bits;
    }
}

__global__ void __launch_bounds__(kBlock) k_logical_not(const int8_t* __restrict__ in,
                                                        const uint64_t* __restrict__ nm,
                                                        int8_t* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        out[r] = (int8_t)(!in[r] && !(nm && null_bit(nm, r)));
}

template <typename T>
T scalar_of(const void* p) {
    T v;
    memcpy(&v, p, sizeof(T));
    return v;
}

template <typename L, typename R>
int launch_typed(bool bits, int32_t op, int32_t lkind, const void* lhs, int32_t rkind,
                 const void* rhs, int64_t n, const uint64_t* nm, void* out, hipStream_t s) {
    const bool lcol = lkind == QK_COLUMN, rcol = rkind == QK_COLUMN;
    const L* a = lcol ? (const L*)lhs : nullptr;
    const R* b = rcol ? (const R*)rhs : nullptr;
    L ac = lcol ? L{} : scalar_of<L>(lhs);
    R bc = rcol ? R{} : scalar_of<R>(rhs);
    if (!lcol && !rcol) {
        // const-const: one value for the whole block (GPUBinary.cuh:575)
        const int v = cmp_op(op, ac, bc) ? 1 : 0;
        if (bits) {
            const int64_t words = (n + 63) / 64;
            QK_HIP(hipMemsetAsync(out, v ? 0xff : 0, (size_t)words * 8, s));
            if (v && (n & 63)) {
                uint64_t last = (1ull << (n & 63)) - 1;
                QK_HIP(hipMemcpyAsync((uint64_t*)out + words - 1, &last, 8, hipMemcpyHostToDevice, s));
                QK_HIP(hipStreamSynchronize(s));
            }
        } else {
            QK_HIP(hipMemsetAsync(out, v, (size_t)n, s));
        }
        return QK_SUCCESS;
    }
    if (bits) {
        int grid = grid_for((n + 63) / 64, kBlock / 64);
        k_filter_bits<L, R><<<grid, kBlock, 0, s>>>(op, a, ac, lcol, b, bc, rcol, nm,
                                                    (uint64_t*)out, n);
    } else if (((uintptr_t)out & 3u) == 0) {
        int grid = grid_for((n + 3) / 4, kBlock);
        k_filter_i8<L, R><<<grid, kBlock, 0, s>>>(op, a, ac, lcol, b, bc, rcol, nm, (int8_t*)out, n);
    } else {
        int grid = grid_for(n, kBlock);
        k_filter_i8_unaligned<L, R><<<grid, kBlock, 0, s>>>(op, a, ac, lcol, b, bc, rcol, nm,
                                                            (int8_t*)out, n);
    }
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

template <typename L>
int launch_rhs(bool bits, int32_t op, int32_t lkind, const void* lhs, int32_t rtype, int32_t rkind,
               const void* rhs, int64_t n, const uint64_t* nm, void* out, hipStream_t s) {
    switch (rtype) {
    case QK_INT8: return launch_typed<L, int8_t>(bits, op, lkind, lhs, rkind, rhs, n, nm, out, s);
    case QK_INT: return launch_typed<L, int32_t>(bits, op, lkind, lhs, rkind, rhs, n, nm, out, s);
    case QK_LONG: return launch_typed<L, int64_t>(bits, op, lkind, lhs, rkind, rhs, n, nm, out, s);
    case QK_FLOAT: return launch_typed<L, float>(bits, op, lkind, lhs, rkind, rhs, n, nm, out, s);
    case QK_DOUBLE: return launch_typed<L, double>(bits, op, lkind, lhs, rkind, rhs, n, nm, out, s);
    default: return fail(QK_ERROR, "filter: unsupported right operand type %d", rtype);
    }
}

int filter_common(bool bits, int32_t op, int32_t ltype, int32_t lkind, const void* lhs,
                  int32_t rtype, int32_t rkind, const void* rhs, int32_t n,
                  const uint64_t* nm, void* out, void* stream) {
    QK_TRY(ensure_init());
    if (op < QK_GREATER || op > QK_LOGICAL_OR) return fail(QK_ERROR, "filter: unknown operator %d", op);
    if ((lkind != QK_COLUMN && lkind != QK_CONST) || (rkind != QK_COLUMN && rkind != QK_CONST))
        return fail(QK_ERROR, "filter: operand kind must be QK_COLUMN or QK_CONST");
    if (!lhs || !rhs) return fail(QK_ERROR, "filter: null operand");
    if (n <= 0) return fail(QK_ERROR, "filter: Data Element Count must be > 0");  // Context.cu:111-115
    if (!out) return fail(QK_ERROR, "filter: null output");
    hipStream_t s = as_stream(stream);
    switch (ltype) {
    case QK_INT8: return launch_rhs<int8_t>(bits, op, lkind, lhs, rtype, rkind, rhs, n, nm, out, s);
    case QK_INT: return launch_rhs<int32_t>(bits, op, lkind, lhs, rtype, rkind, rhs, n, nm, out, s);
    case QK_LONG: return launch_rhs<int64_t>(bits, op, lkind, lhs, rtype, rkind, rhs, n, nm, out, s);
    case QK_FLOAT: return launch_rhs<float>(bits, op, lkind, lhs, rtype, rkind, rhs, n, nm, out, s);
    case QK_DOUBLE: return launch_rhs<double>(bits, op, lkind, lhs, rtype, rkind, rhs, n, nm, out, s);
    default: return fail(QK_ERROR, "filter: unsupported left operand type %d", ltype);
    }
}

}  // namespace

int check_preds(int32_t npreds, const qk_predicate* preds, PredSet* out) {
    if (npreds < 0 || npreds > kMaxPreds)
        return fail(QK_ERROR, "at most %d predicates are supported (got %d)", kMaxPreds, npreds);
    if (npreds > 0 && !preds) return fail(QK_ERROR, "null predicate array");
    out->n = npreds;
    for (int i = 0; i < npreds; i++) {
        const qk_predicate& p = preds[i];
        if (p.op < QK_GREATER || p.op > QK_LOGICAL_OR)
            return fail(QK_ERROR, "predicate %d: unknown operator %d", i, p.op);
        if (!type_size(p.type) || !type_size(p.rhs_type))
            return fail(QK_ERROR, "predicate %d: unsupported operand type", i);
        if (!p.column) return fail(QK_ERROR, "predicate %d: null column", i);
        if (p.rhs_kind != QK_CONST && p.rhs_kind != QK_COLUMN)
            return fail(QK_ERROR, "predicate %d: rhs_kind must be QK_CONST or QK_COLUMN", i);
        if (p.rhs_kind == QK_COLUMN && !p.rhs_column)
            return fail(QK_ERROR, "predicate %d: null right column", i);
        out->p[i] = p;
    }
    return QK_SUCCESS;
}

namespace {

// One bound of one domain while folding: value and whether it is exclusive.
template <typename T>
struct Bound {
    T v;
    bool strict;
};

template <typename T>
void tighten(int op, T k, Bound<T>& lo, Bound<T>& hi) {
    const bool gt = op == QK_GREATER, ge = op == QK_GREATER_EQUAL || op == QK_EQUAL;
    const bool lt = op == QK_LESS, le = op == QK_LESS_EQUAL || op == QK_EQUAL;
    if ((gt || ge) && (k > lo.v || (k == lo.v && gt))) lo = {k, gt};
    if ((lt || le) && (k < hi.v || (k == hi.v && lt))) hi = {k, lt};
}

struct Fold {
    int32_t doms = 0;
    Bound<int64_t> lo_i{INT64_MIN, false}, hi_i{INT64_MAX, false};
    Bound<float> lo_f{-INFINITY, false}, hi_f{INFINITY, false};
    Bound<double> lo_d{-INFINITY, false}, hi_d{INFINITY, false};
};

// One comparison `v op c` with v of type VT and c of type R, in their common type.
template <typename VT, typename R>
void fold_one(int op, R c, Fold& f) {
    using D = typename std::common_type<VT, R>::type;
    const D k = (D)c;
    if constexpr (std::is_integral<D>::value) {
        f.doms |= 1;
        tighten<int64_t>(op, (int64_t)k, f.lo_i, f.hi_i);
    } else if constexpr (std::is_same<D, float>::value) {
        f.doms |= 2;
        tighten<float>(op, k, f.lo_f, f.hi_f);
    } else {
        f.doms |= 4;
        tighten<double>(op, (double)k, f.lo_d, f.hi_d);
    }
}

template <typename VT>
void fold_typed(const qk_predicate& p, Fold& f) {
    switch (p.rhs_type) {
    case QK_INT8: fold_one<VT, int8_t>(p.op, (int8_t)p.rhs_int, f); break;
    case QK_INT: fold_one<VT, int32_t>(p.op, (int32_t)p.rhs_int, f); break;
    case QK_LONG: fold_one<VT, int64_t>(p.op, (int64_t)p.rhs_int, f); break;
    case QK_FLOAT: fold_one<VT, float>(p.op, (float)p.rhs_real, f); break;
    default: fold_one<VT, double>(p.op, p.rhs_real, f); break;
    }
}

// A floating bound as the pair the kernels test (the unused member NaN).
template <typename T>
void split(const Bound<T>& b, T& x, T& i) {
    x = b.strict ? b.v : (T)NAN;
    i = b.strict ? (T)NAN : b.v;
}

}  // namespace

namespace {

// The interval the kernels test, from the folded bounds.
ValueRange range_of(const Fold& f) {
    ValueRange vr{};
    vr.doms = f.doms;
    // integer bounds inclusive; a strict bound at the end of the range empties the interval
    vr.lo_i = f.lo_i.v;
    vr.hi_i = f.hi_i.v;
    if (f.lo_i.strict) {
        if (vr.lo_i == INT64_MAX) { vr.lo_i = INT64_MAX; vr.hi_i = INT64_MIN; } else vr.lo_i++;
    }
    if (f.hi_i.strict) {
        if (vr.hi_i == INT64_MIN) { vr.lo_i = INT64_MAX; vr.hi_i = INT64_MIN; } else vr.hi_i--;
    }
    split(f.lo_f, vr.lo_fx, vr.lo_fi);
    split(f.hi_f, vr.hi_fx, vr.hi_fi);
    split(f.lo_d, vr.lo_dx, vr.lo_di);
    split(f.hi_d, vr.hi_dx, vr.hi_di);
    return vr;
}

bool foldable_op(int32_t op) {
    return op == QK_GREATER || op == QK_GREATER_EQUAL || op == QK_LESS || op == QK_LESS_EQUAL || op == QK_EQUAL;
}

}  // namespace

bool fold_value_range(PredSet* ps, uint32_t vpred, int32_t vtype) {
    if (ps->n == 0) return false;
    Fold f;
    for (int i = 0; i < ps->n; i++) {
        const qk_predicate& p = ps->p[i];
        if (!((vpred >> i) & 1u) || p.rhs_kind != QK_CONST || !foldable_op(p.op)) return false;
        switch (vtype) {
        case QK_INT: fold_typed<int32_t>(p, f); break;
        case QK_LONG: fold_typed<int64_t>(p, f); break;
        case QK_FLOAT: fold_typed<float>(p, f); break;
        case QK_DOUBLE: fold_typed<double>(p, f); break;
        default: return false;
        }
    }
    ps->vr = range_of(f);
    return true;
}

bool fold_column_range(PredSet* ps, uint32_t vpred) {
    if (ps->n == 0 || vpred != 0) return false;
    const qk_predicate& p0 = ps->p[0];
    if ((p0.type != QK_INT && p0.type != QK_LONG) || !p0.column) return false;
    Fold f;
    for (int i = 0; i < ps->n; i++) {
        const qk_predicate& p = ps->p[i];
        if (p.column != p0.column || p.type != p0.type || p.null_mask != p0.null_mask || p.rhs_kind != QK_CONST ||
            !foldable_op(p.op))
            return false;
        if (p0.type == QK_INT) fold_typed<int32_t>(p, f); else fold_typed<int64_t>(p, f);
    }
    ps->vr = range_of(f);
    ps->rcol = p0.column;
    ps->rnull = p0.null_mask;
    ps->rwide = p0.type == QK_LONG ? 1 : 0;
    return true;
}

}  // namespace qk

using namespace qk;

extern "C" {

int qk_filter(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
              int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask,
              int8_t* out_mask, void* stream) {
    return filter_common(false, op, ltype, lkind, lhs, rtype, rkind, rhs, n, null_mask, out_mask,
                         stream);
}

int qk_filter_bits(int32_t op, int32_t ltype, int32_t lkind, const void* lhs, int32_t rtype,
                   int32_t rkind, const void* rhs, int32_t n, const uint64_t* null_mask,
                   uint64_t* out_bits, void* stream) {
    return filter_common(true, op, ltype, lkind, lhs, rtype, rkind, rhs, n, null_mask, out_bits,
                         stream);
}

int qk_logical_not(const int8_t* in, int32_t n, const uint64_t* null_mask, int8_t* out,
                   void* stream) {
    QK_TRY(ensure_init());
    if (n <= 0) return fail(QK_ERROR, "logical_not: Data Element Count must be > 0");
    if (!in || !out) return fail(QK_ERROR, "logical_not: null buffer");
    int grid = grid_for(n, kBlock);
    k_logical_not<<<grid, kBlock, 0, as_stream(stream)>>>(in, null_mask, out, n);
    QK_HIP(hipGetLastError());
    return QK_SUCCESS;
}

}  // extern "C"
