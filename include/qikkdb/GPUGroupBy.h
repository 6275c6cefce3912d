// GPUGroupBy.h - GROUP BY tables of the C++ drop-in layer, with the call shapes of the
// reference's GPUGroupBy<AGG,O,K,V> (QueryEngine/GPUCore/GPUGroupBy.cuh:276-591) and its
// composite-key form GPUGroupBy<AGG,O,std::vector<void*>,V> (GPUGroupByMultiKey.cuh:282-923).
// Each table forwards to one qk_groupby handle of the C ABI.
//
// Not provided: the pre-seeded constructor (GPUGroupBy.cuh:367-372) and the
// GetResults(..., tables, ...) / ReconstructRawNumbers overloads (:433-790), which exist for
// the reference's host-staged merge of one table per GPU; this build runs on one GPU.
#pragma once

#include <cstdint>
#include <type_traits>
#include <vector>

#include "AggregationFunctions.h"
#include "DataType.h"
#include "GPUBinary.h"
#include "GPUError.h"
#include "GPUMemory.h"
#include "NullValues.h"

// Type-erased handle (IGroupBy.h:9-20).
class IGroupBy {
public:
    virtual ~IGroupBy() = default;
};

namespace qikkdb {

// The typing rules of GPUGroupBy.cuh:522-529 and :563-567: COUNT gives int64, every other
// aggregate gives the value type (this build also accepts a widened SUM: int32 -> int64,
// float -> double).
template <typename AGG, typename O, typename V>
constexpr void check_output_type() {
    static_assert(!std::is_same<AGG, AggregationFunctions::count>::value || std::is_same<O, int64_t>::value,
                  "GPUGroupBy: COUNT outputs int64_t");
}

class Table {
    qk_groupby h_ = nullptr;

public:
    Table(int32_t nkeys, const int32_t* key_types, int32_t value_type, int32_t agg, int32_t out_type,
          int32_t maxHashCount, int32_t keySizeMultiplier) {
        const int32_t out_types[1] = {out_type};
        check(qk_gb_create(&h_, nkeys, key_types, value_type, 1, &agg, out_types, maxHashCount,
                           keySizeMultiplier, 0));
    }
    Table(const Table&) = delete;
    Table& operator=(const Table&) = delete;
    ~Table() {
        if (h_) qk_gb_destroy(h_);
    }
    qk_groupby get() const { return h_; }
};

}  // namespace qikkdb

template <typename AGG, typename O, typename K, typename V>
class GPUGroupBy : public IGroupBy {
    static constexpr bool USE_VALUES = std::is_same<AGG, AggregationFunctions::min>::value ||
                                       std::is_same<AGG, AggregationFunctions::max>::value ||
                                       std::is_same<AGG, AggregationFunctions::sum>::value ||
                                       std::is_same<AGG, AggregationFunctions::avg>::value;
    qikkdb::Table table_;

    static const int32_t* key_type() {
        static const int32_t t[1] = {qikkdb::TypeCode<K>::value};
        return t;
    }

public:
    GPUGroupBy(const int32_t maxHashCount, const int32_t keySizeMultiplier)
    : table_(1, key_type(), qikkdb::TypeCode<V>::value, AGG::code, qikkdb::TypeCode<O>::value, maxHashCount,
             keySizeMultiplier) {
        qikkdb::check_output_type<AGG, O, V>();
    }

    // Aggregate one block; NULL masks may be nullptr.  Waits for the kernel and throws
    // query_engine_error(GPU_HASH_TABLE_FULL) when the table filled up.
    void ProcessBlock(K* inKeys, V* inValues, int32_t dataElementCount, nullmask_t* inKeysNullMask = nullptr,
                      nullmask_t* inValuesNullMask = nullptr) {
        if (dataElementCount <= 0) return;
        const void* keys[1] = {inKeys};
        const uint64_t* knulls[1] = {inKeysNullMask};
        qikkdb::check(qk_gb_process_block(table_.get(), keys, inKeysNullMask ? knulls : nullptr,
                                          USE_VALUES ? inValues : nullptr, inValuesNullMask, dataElementCount,
                                          nullptr));
    }

    int32_t GetMaxHashCount() {
        int32_t v = 0;
        qikkdb::check(qk_gb_get_max_hash_count(table_.get(), &v));
        return v;
    }

    // One row per group; outputs are device buffers the caller frees with GPUMemory::free.
    void GetResults(K** outKeys, O** outValues, int32_t* outDataElementCount, nullmask_t** outKeysNullMask = nullptr,
                    nullmask_t** outValuesNullMask = nullptr) {
        void* keys[1] = {nullptr};
        uint64_t* knulls[1] = {nullptr};
        void* vals[1] = {nullptr};
        uint64_t* vnulls[1] = {nullptr};
        qikkdb::check(qk_gb_get_results(table_.get(), keys, outKeysNullMask ? knulls : nullptr, vals,
                                        outValuesNullMask ? vnulls : nullptr, outDataElementCount, nullptr));
        *outKeys = static_cast<K*>(keys[0]);
        *outValues = static_cast<O*>(vals[0]);
        if (outKeysNullMask) *outKeysNullMask = knulls[0];
        if (outValuesNullMask) *outValuesNullMask = vnulls[0];
    }
};

template <typename AGG, typename O, typename V>
class GPUGroupBy<AGG, O, std::vector<void*>, V> : public IGroupBy {
    static constexpr bool USE_VALUES = std::is_same<AGG, AggregationFunctions::min>::value ||
                                       std::is_same<AGG, AggregationFunctions::max>::value ||
                                       std::is_same<AGG, AggregationFunctions::sum>::value ||
                                       std::is_same<AGG, AggregationFunctions::avg>::value;
    std::vector<DataType> keyTypes_;
    std::vector<int32_t> codes_;
    qikkdb::Table table_;

    static std::vector<int32_t> codes_of(const std::vector<DataType>& t) {
        return std::vector<int32_t>(t.begin(), t.end());
    }

public:
    GPUGroupBy(const int32_t maxHashCount, const int32_t keySizeMultiplier, std::vector<DataType>& keyTypes)
    : keyTypes_(keyTypes), codes_(codes_of(keyTypes)),
      table_((int32_t)codes_.size(), codes_.data(), qikkdb::TypeCode<V>::value, AGG::code,
             qikkdb::TypeCode<O>::value, maxHashCount, keySizeMultiplier) {
        qikkdb::check_output_type<AGG, O, V>();
    }

    void ProcessBlock(std::vector<void*> inKeysVector, std::vector<nullmask_t*> inKeysNullMaskVector, V* inValues,
                      int32_t dataElementCount, nullmask_t* inValuesNullMask = nullptr) {
        if (dataElementCount <= 0) return;
        if (inKeysVector.size() != keyTypes_.size())
            throw query_engine_error(GPU_EXTENSION_ERROR, "GPUGroupBy: wrong number of key columns");
        std::vector<const void*> keys(inKeysVector.begin(), inKeysVector.end());
        std::vector<const uint64_t*> knulls(keys.size(), nullptr);
        for (size_t i = 0; i < knulls.size() && i < inKeysNullMaskVector.size(); i++) knulls[i] = inKeysNullMaskVector[i];
        qikkdb::check(qk_gb_process_block(table_.get(), keys.data(), knulls.data(), USE_VALUES ? inValues : nullptr,
                                          inValuesNullMask, dataElementCount, nullptr));
    }

    int32_t GetMaxHashCount() {
        int32_t v = 0;
        qikkdb::check(qk_gb_get_max_hash_count(table_.get(), &v));
        return v;
    }

    void GetResults(std::vector<void*>* outKeysVector, O** outValues, int32_t* outDataElementCount,
                    std::vector<nullmask_t*>* outKeysNullMasksVector = nullptr, nullmask_t** outValuesNullMask = nullptr) {
        const size_t nk = keyTypes_.size();
        std::vector<void*> keys(nk, nullptr);
        std::vector<uint64_t*> knulls(nk, nullptr);
        void* vals[1] = {nullptr};
        uint64_t* vnulls[1] = {nullptr};
        qikkdb::check(qk_gb_get_results(table_.get(), keys.data(), outKeysNullMasksVector ? knulls.data() : nullptr,
                                        vals, outValuesNullMask ? vnulls : nullptr, outDataElementCount, nullptr));
        *outKeysVector = keys;
        if (outKeysNullMasksVector) outKeysNullMasksVector->assign(knulls.begin(), knulls.end());
        *outValues = static_cast<O*>(vals[0]);
        if (outValuesNullMask) *outValuesNullMask = vnulls[0];
    }
};
