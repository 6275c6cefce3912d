// GPUBinary.h - the filter form of GPUBinary (QueryEngine/GPUCore/GPUBinary.cuh:383-401 and
// the filter specialisation at :554-579 of the reference): out[i] = OP(A, B) as an int8 0/1
// mask, where A and B are each a device column (T*) or a scalar (T), with C++ comparison
// typing and NULL rows (nullBitMask, the OR of the operands' masks) giving 0.
#pragma once

#include <cstdint>
#include <type_traits>

#include "FilterConditions.h"
#include "GPUError.h"
#include "NullValues.h"

namespace qikkdb {

template <typename T>
struct TypeCode;
template <> struct TypeCode<int8_t> { static constexpr int32_t value = QK_INT8; };
template <> struct TypeCode<int32_t> { static constexpr int32_t value = QK_INT; };
template <> struct TypeCode<int64_t> { static constexpr int32_t value = QK_LONG; };
template <> struct TypeCode<float> { static constexpr int32_t value = QK_FLOAT; };
template <> struct TypeCode<double> { static constexpr int32_t value = QK_DOUBLE; };

// An operand: a device column when P is a pointer, else a host scalar passed by address.
template <typename P>
struct Operand {
    using elem = typename std::remove_cv<typename std::remove_pointer<P>::type>::type;
    static constexpr bool column = std::is_pointer<P>::value;
    static constexpr int32_t type = TypeCode<elem>::value;
    static constexpr int32_t kind = column ? QK_COLUMN : QK_CONST;
    static const void* ptr(const P& p) {
        if constexpr (column) return static_cast<const void*>(p);
        else return static_cast<const void*>(&p);
    }
};

}  // namespace qikkdb

template <typename OP, typename OUT, typename L, typename R>
class GPUBinary {
public:
    static_assert(std::is_same<OUT, int8_t>::value,
                  "qikkdb_amd provides the filter (int8 mask) form of GPUBinary");
    static void Binary(OUT* output, L ACol, R BCol, int32_t dataElementCount, nullmask_t* nullBitMask = nullptr) {
        using A = qikkdb::Operand<L>;
        using B = qikkdb::Operand<R>;
        qikkdb::check(qk_filter(OP::code, A::type, A::kind, A::ptr(ACol), B::type, B::kind, B::ptr(BCol),
                                dataElementCount, nullBitMask, output, nullptr));
    }
};
