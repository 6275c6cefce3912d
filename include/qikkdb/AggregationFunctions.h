// AggregationFunctions.h - the aggregate tags of GPUGroupBy<AGG,...>, named as in
// QueryEngine/GPUCore/AggregationFunctions.cuh:14-367 of the reference.  Here they only
// select the table's aggregate; the arithmetic lives in the HIP kernels.
#pragma once

#include <cstdint>

#include "../qikkdb_amd.h"

namespace AggregationFunctions {
struct min { static constexpr int32_t code = QK_AGG_MIN; };
struct max { static constexpr int32_t code = QK_AGG_MAX; };
struct sum { static constexpr int32_t code = QK_AGG_SUM; };
struct avg { static constexpr int32_t code = QK_AGG_AVG; };
struct count { static constexpr int32_t code = QK_AGG_COUNT; };
struct none { static constexpr int32_t code = QK_AGG_NONE; };
}  // namespace AggregationFunctions
