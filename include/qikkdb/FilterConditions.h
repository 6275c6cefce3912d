// FilterConditions.h - the comparison tags of GPUBinary<OP,int8_t,L,R>, named as in
// QueryEngine/GPUCore/GPUFilterConditions.cuh:6-180 of the reference.
#pragma once

#include <cstdint>

#include "../qikkdb_amd.h"

namespace FilterConditions {
struct greater { static constexpr int32_t code = QK_GREATER; };
struct greaterEqual { static constexpr int32_t code = QK_GREATER_EQUAL; };
struct less { static constexpr int32_t code = QK_LESS; };
struct lessEqual { static constexpr int32_t code = QK_LESS_EQUAL; };
struct equal { static constexpr int32_t code = QK_EQUAL; };
struct notEqual { static constexpr int32_t code = QK_NOT_EQUAL; };
struct logicalAnd { static constexpr int32_t code = QK_LOGICAL_AND; };
struct logicalOr { static constexpr int32_t code = QK_LOGICAL_OR; };
}  // namespace FilterConditions
