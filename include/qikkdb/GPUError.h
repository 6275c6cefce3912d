// GPUError.h - error types of the C++ drop-in layer.
//
// Mirrors QueryEngine/GPUError.h:17-75 of the reference: QueryEngineErrorType (same values),
// gpu_error and query_engine_error, and CheckQueryEngineError.  Every wrapper in this folder
// turns a status of the C ABI (include/qikkdb_amd.h) into these exceptions:
//   QK_HASH_TABLE_FULL / QK_DIVISION_BY_ZERO / QK_INTEGER_OVERFLOW / QK_UNKNOWN_AGG_FUN / QK_ERROR
//       -> query_engine_error with the matching QueryEngineErrorType
//   QK_OUT_OF_MEMORY -> std::out_of_range (GPUMemory.cuh:95-107 rethrows that after eviction)
//   QK_RUNTIME_ERROR -> gpu_error (the reference aborts the process, GPUError.cpp:37)
#pragma once

#include <stdexcept>
#include <string>

#include "../qikkdb_amd.h"

enum QueryEngineErrorType {
    GPU_EXTENSION_SUCCESS = 0,
    GPU_EXTENSION_ERROR,
    GPU_DIVISION_BY_ZERO_ERROR,
    GPU_INTEGER_OVERFLOW_ERROR,
    GPU_HASH_TABLE_FULL,
    GPU_UNKNOWN_AGG_FUN,
    GPU_NOT_FOUND_ERROR,
    GPU_MEMORY_MAPPING_NOT_SUPPORTED_ERROR,
    GPU_DRIVER_NOT_FOUND_EXCEPTION
};

class gpu_error : public std::runtime_error {
public:
    explicit gpu_error(const std::string& message) : std::runtime_error(message) {}
};

class query_engine_error : public gpu_error {
    QueryEngineErrorType queryEngineErrorType_;

public:
    query_engine_error(QueryEngineErrorType type, const std::string& message)
    : gpu_error("Query engine error " + std::to_string(static_cast<int>(type)) + ": " + message),
      queryEngineErrorType_(type) {}
    QueryEngineErrorType GetQueryEngineError() const { return queryEngineErrorType_; }
};

inline void CheckQueryEngineError(const QueryEngineErrorType errorType, const std::string& message = std::string()) {
    if (errorType != GPU_EXTENSION_SUCCESS) throw query_engine_error(errorType, message);
}

namespace qikkdb {

// Status of a qk_* call -> exception (see the table at the top of this file).
inline void check(int status) {
    if (status == QK_SUCCESS) return;
    const std::string msg = qk_last_error_message();
    switch (status) {
    case QK_HASH_TABLE_FULL: throw query_engine_error(GPU_HASH_TABLE_FULL, msg);
    case QK_DIVISION_BY_ZERO: throw query_engine_error(GPU_DIVISION_BY_ZERO_ERROR, msg);
    case QK_INTEGER_OVERFLOW: throw query_engine_error(GPU_INTEGER_OVERFLOW_ERROR, msg);
    case QK_UNKNOWN_AGG_FUN: throw query_engine_error(GPU_UNKNOWN_AGG_FUN, msg);
    case QK_OUT_OF_MEMORY: throw std::out_of_range(msg);
    case QK_RUNTIME_ERROR: throw gpu_error(msg);
    default: throw query_engine_error(GPU_EXTENSION_ERROR, msg);
    }
}

}  // namespace qikkdb
