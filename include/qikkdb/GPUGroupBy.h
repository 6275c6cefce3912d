// GPUGroupBy.h - GROUP BY tables of the C++ drop-in layer, with the call shapes of the
// reference's GPUGroupBy<AGG,O,K,V> (QueryEngine/GPUCore/GPUGroupBy.cuh:276-591) and its
// composite-key form GPUGroupBy<AGG,O,std::vector<void*>,V> (GPUGroupByMultiKey.cuh:282-923).
// Each table forwards to one qk_groupby handle of the C ABI.
//
// GetResults(..., tables, ...) (GPUGroupBy.cuh:593-790, GPUGroupByMultiKey.cuh:924-1324) is the
// only GetResults the reference's dispatcher calls (GpuSqlDispatcherAggregationFunctions.h:141,
// :203, :303).  With one table it is the single-table GetResults, as in the reference; with
// several tables (one per dispatcher thread) the others are folded into this one on the device
// (qk_gb_merge) instead of being staged through the host.  Tables on different GPUs would need
// peer access, which this one-GPU build does not set up: that case throws.
// The composite form has no ReconstructRawNumbers (its reference version fills host key
// columns of every type, strings included, for the host-staged merge only).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
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

// Fold the dispatcher's other tables into `self` (the multi-table GetResults).  As in the
// reference, an empty vector is an error and a vector whose second entry is null holds one table.
template <typename T>
void merge_tables(T* self, qk_groupby self_h, std::vector<std::unique_ptr<IGroupBy>>& tables) {
    if (tables.size() == 1 || (tables.size() > 1 && tables[1].get() == nullptr)) return;
    for (auto& t : tables) {
        if (!t || t.get() == self) continue;
        T* other = dynamic_cast<T*>(t.get());
        if (!other)
            throw query_engine_error(GPU_EXTENSION_ERROR, "GetResults: the tables are not of the same GPUGroupBy type");
        check(qk_gb_merge(self_h, other->handle(), nullptr));
    }
}

// The typing rules of GPUGroupBy.cuh:522-529 and :563-567: COUNT gives int64, every other
// aggregate gives the value type (the library also accepts a widened SUM: int32 -> int64,
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

    // A table holding the maxHashCount keys of the device buffer `keys` (keys[0] stands for
    // the NULL key when firstKeyIsNull) as groups from the start; GetResults returns them first,
    // in the order of `keys` (GPUGroupBy.cuh:367-372, where they fill consecutive slots).
    GPUGroupBy(const int32_t maxHashCount, K* keys, bool firstKeyIsNull)
    : table_(1, key_type(), qikkdb::TypeCode<V>::value, AGG::code, qikkdb::TypeCode<O>::value,
             maxHashCount - (firstKeyIsNull ? 1 : 0) > 0 ? maxHashCount - (firstKeyIsNull ? 1 : 0) : 1, 1) {
        qikkdb::check_output_type<AGG, O, V>();
        qikkdb::check(qk_gb_seed_keys(table_.get(), keys, maxHashCount, firstKeyIsNull ? 1 : 0, nullptr));
    }

    qk_groupby handle() const { return table_.get(); }

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

    // The dispatcher's GetResults over the tables of all its threads (GPUGroupBy.cuh:593-790).
    void GetResults(K** outKeys, O** outValues, int32_t* outDataElementCount,
                    std::vector<std::unique_ptr<IGroupBy>>& tables, nullmask_t** outKeysNullMask = nullptr,
                    nullmask_t** outValuesNullMask = nullptr) {
        if (tables.size() <= 0) throw std::invalid_argument("Number of tables have to be at least 1.");
        qikkdb::merge_tables(this, table_.get(), tables);
        GetResults(outKeys, outValues, outDataElementCount, outKeysNullMask, outValuesNullMask);
    }

    // Raw fields before finalisation, copied to HOST buffers of at least GetMaxHashCount()
    // entries (GPUGroupBy.cuh:433-471): keys and one NULL byte per key; for MIN/MAX/SUM/AVG the
    // merged values (the sum for AVG) and one NULL byte per value; for AVG/COUNT the row counts.
    void ReconstructRawNumbers(K* keys, nullarray_t* keysNullArray, V* values, nullarray_t* valuesNullArray,
                               int64_t* occurrences, int32_t* elementCount) {
        constexpr bool OCC = std::is_same<AGG, AggregationFunctions::avg>::value ||
                             std::is_same<AGG, AggregationFunctions::count>::value;
        const int32_t vagg = std::is_same<AGG, AggregationFunctions::avg>::value ? AggregationFunctions::sum::code
                                                                                  : AGG::code;
        int32_t aggs[2] = {USE_VALUES ? vagg : QK_AGG_NONE, QK_AGG_COUNT};
        int32_t otypes[2] = {qikkdb::TypeCode<V>::value, QK_LONG};
        void* k[1] = {nullptr};
        uint64_t* kn[1] = {nullptr};
        void* v[2] = {nullptr, nullptr};
        uint64_t* vn[2] = {nullptr, nullptr};
        int32_t g = 0;
        qikkdb::check(qk_gb_get_results_aggs(table_.get(), 2, aggs, otypes, k, kn, v, vn, &g, nullptr));
        cuda_ptr_set owned{{k[0], kn[0], v[0], v[1], vn[0], vn[1]}};
        *elementCount = g;
        if (g == 0) return;
        std::vector<uint64_t> bits(NullValues::GetNullBitMaskSize(g));
        auto unpack = [&](const uint64_t* dbits, nullarray_t* out) {
            GPUMemory::copyDeviceToHost(bits.data(), dbits, bits.size());
            for (int32_t i = 0; i < g; i++) out[i] = (nullarray_t)((bits[i >> 6] >> (i & 63)) & 1ull);
        };
        GPUMemory::copyDeviceToHost(keys, static_cast<K*>(k[0]), g);
        if (keysNullArray) unpack(kn[0], keysNullArray);
        if (USE_VALUES) {
            if (values) GPUMemory::copyDeviceToHost(values, static_cast<V*>(v[0]), g);
            if (valuesNullArray) unpack(vn[0], valuesNullArray);
        }
        if (OCC && occurrences) GPUMemory::copyDeviceToHost(occurrences, static_cast<int64_t*>(v[1]), g);
    }

private:
    // frees the device buffers of one results read
    struct cuda_ptr_set {
        std::vector<void*> p;
        ~cuda_ptr_set() {
            for (void* x : p)
                if (x) qk_free(x);
        }
    };
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

    qk_groupby handle() const { return table_.get(); }

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

    // The dispatcher's GetResults over the tables of all its threads (GPUGroupByMultiKey.cuh:924-1324).
    void GetResults(std::vector<void*>* outKeysVector, O** outValues, int32_t* outDataElementCount,
                    std::vector<std::unique_ptr<IGroupBy>>& tables,
                    std::vector<nullmask_t*>* outKeysNullMasksVector = nullptr, nullmask_t** outValuesNullMask = nullptr) {
        if (tables.size() <= 0) CheckQueryEngineError(GPU_EXTENSION_ERROR, "Number of tables have to be at least 1.");
        qikkdb::merge_tables(this, table_.get(), tables);
        GetResults(outKeysVector, outValues, outDataElementCount, outKeysNullMasksVector, outValuesNullMask);
    }
};

// GROUP BY over one string key column (GPUGroupByString.cuh:169-785): keys arrive and leave as
// GPUString columns; every other rule is the numeric table's.
template <typename AGG, typename O, typename V>
class GPUGroupBy<AGG, O, std::string, V> : public IGroupBy {
    static constexpr bool USE_VALUES = std::is_same<AGG, AggregationFunctions::min>::value ||
                                       std::is_same<AGG, AggregationFunctions::max>::value ||
                                       std::is_same<AGG, AggregationFunctions::sum>::value ||
                                       std::is_same<AGG, AggregationFunctions::avg>::value;
    qk_gbs h_ = nullptr;

public:
    GPUGroupBy(const int32_t maxHashCount, const int32_t keySizeMultiplier) {
        qikkdb::check_output_type<AGG, O, V>();
        const int32_t agg = AGG::code, ot = qikkdb::TypeCode<O>::value;
        qikkdb::check(qk_gbs_create(&h_, qikkdb::TypeCode<V>::value, 1, &agg, &ot, maxHashCount, keySizeMultiplier, 0));
    }
    GPUGroupBy(const GPUGroupBy&) = delete;
    GPUGroupBy& operator=(const GPUGroupBy&) = delete;
    ~GPUGroupBy() {
        if (h_) qk_gbs_destroy(h_);
    }

    void ProcessBlock(GPUMemory::GPUString inKeys, V* inValues, int32_t dataElementCount,
                      nullmask_t* inKeysNullMask = nullptr, nullmask_t* inValuesNullMask = nullptr) {
        if (dataElementCount <= 0) return;
        qikkdb::check(qk_gbs_process_block(h_, inKeys.allChars, inKeys.stringIndices, inKeysNullMask,
                                           USE_VALUES ? inValues : nullptr, inValuesNullMask, dataElementCount,
                                           nullptr));
    }

    void GetResults(GPUMemory::GPUString* outKeys, O** outValues, int32_t* outDataElementCount,
                    nullmask_t** outKeysNullMask = nullptr, nullmask_t** outValuesNullMask = nullptr) {
        void* vals[1] = {nullptr};
        uint64_t* vnulls[1] = {nullptr};
        uint64_t* kn = nullptr;
        qikkdb::check(qk_gbs_get_results(h_, &outKeys->allChars, &outKeys->stringIndices, &kn, vals,
                                         outValuesNullMask ? vnulls : nullptr, outDataElementCount, nullptr));
        *outValues = static_cast<O*>(vals[0]);
        if (outKeysNullMask) *outKeysNullMask = kn;
        else if (kn) qk_free(kn);
        if (outValuesNullMask) *outValuesNullMask = vnulls[0];
    }

    // The dispatcher's GetResults over its tables; one table in this one-GPU build.
    void GetResults(GPUMemory::GPUString* outKeys, O** outValues, int32_t* outDataElementCount,
                    std::vector<std::unique_ptr<IGroupBy>>& tables, nullmask_t** outKeysNullMask = nullptr,
                    nullmask_t** outValuesNullMask = nullptr) {
        if (tables.size() <= 0) throw std::invalid_argument("Number of tables have to be at least 1.");
        for (auto& t : tables)
            if (t && t.get() != this)
                throw query_engine_error(GPU_EXTENSION_ERROR,
                                         "GetResults: string keys merge one table per GPU; this is a one-GPU build");
        GetResults(outKeys, outValues, outDataElementCount, outKeysNullMask, outValuesNullMask);
    }
};
