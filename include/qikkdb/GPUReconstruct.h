// GPUReconstruct.h - stream compaction of the C++ drop-in layer, with the call shapes of
// QueryEngine/GPUCore/GPUReconstruct.cuh:189-627 in the reference (numeric columns).
// *Keep outputs are allocated here and released by the caller with GPUMemory::free; a
// nullptr mask keeps every row and aliases the input (no ownership passes), and zero kept
// rows give nullptr outputs, as in the reference (GPUReconstruct.cuh:311-338).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "GPUError.h"
#include "GPUMemory.h"
#include "NullValues.h"

// Owning device pointer (cuda_ptr.h:7-68 of the reference).
template <typename T>
class cuda_ptr {
    T* p_ = nullptr;

public:
    cuda_ptr() = default;
    explicit cuda_ptr(T* p) : p_(p) {}
    cuda_ptr(const cuda_ptr&) = delete;
    cuda_ptr& operator=(const cuda_ptr&) = delete;
    cuda_ptr(cuda_ptr&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
    cuda_ptr& operator=(cuda_ptr&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            o.p_ = nullptr;
        }
        return *this;
    }
    ~cuda_ptr() { reset(); }
    T* get() const { return p_; }
    T* release() {
        T* p = p_;
        p_ = nullptr;
        return p;
    }
    void reset() {
        if (p_) qk_free(p_);
        p_ = nullptr;
    }
};

class GPUReconstruct {
public:
    // Compact inCol by inMask and copy the kept rows (and null bits) to host buffers.
    template <typename T>
    static void reconstructCol(T* outData, int32_t* outDataElementCount, T* ACol, int8_t* inMask,
                               int32_t dataElementCount, nullmask_t* outNullMask = nullptr,
                               nullmask_t* nullMask = nullptr) {
        T* dev = nullptr;
        nullmask_t* devNull = nullptr;
        reconstructColKeep(&dev, outDataElementCount, ACol, inMask, dataElementCount,
                           outNullMask ? &devNull : nullptr, nullMask);
        const int32_t n = *outDataElementCount;
        if (n > 0) {
            GPUMemory::copyDeviceToHost(outData, dev, n);
            if (outNullMask && devNull) GPUMemory::copyDeviceToHost(outNullMask, devNull, NullValues::GetNullBitMaskSize(n));
        }
        if (inMask) GPUMemory::free(dev);
        GPUMemory::free(devNull);
    }

    template <typename T>
    static void reconstructColKeep(T** outCol, int32_t* outDataElementCount, T* ACol, int8_t* inMask,
                                   int32_t dataElementCount, nullmask_t** outNullMask = nullptr,
                                   nullmask_t* nullMask = nullptr) {
        void* out = nullptr;
        qikkdb::check(qk_reconstruct_col_keep((int32_t)sizeof(T), ACol, inMask, dataElementCount, nullMask, &out,
                                              outDataElementCount, outNullMask, nullptr));
        *outCol = static_cast<T*>(out);
    }

    // Row indices of the non-zero mask entries, copied to a host buffer.
    template <typename T, typename M>
    static void GenerateIndexes(T* outData, int32_t* outDataElementCount, M* inMask, int32_t dataElementCount) {
        T* dev = nullptr;
        GenerateIndexesKeep(&dev, outDataElementCount, inMask, dataElementCount);
        if (*outDataElementCount > 0) GPUMemory::copyDeviceToHost(outData, dev, *outDataElementCount);
        GPUMemory::free(dev);
    }

    template <typename T, typename M>
    static void GenerateIndexesKeep(T** outData, int32_t* outDataElementCount, M* inMask, int32_t dataElementCount) {
        static_assert(sizeof(M) == 1, "masks are one byte per row");
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "indices are int32 or int64");
        void* out = nullptr;
        qikkdb::check(qk_generate_indexes_keep(reinterpret_cast<const int8_t*>(inMask), dataElementCount,
                                               (int32_t)sizeof(T), &out, outDataElementCount, nullptr));
        *outData = static_cast<T*>(out);
    }

    // Inclusive int32 prefix sum of a byte mask.
    template <typename T, typename M>
    static void PrefixSum(T* prefixSumPointer, M* inMask, int32_t dataElementCount) {
        static_assert(sizeof(T) == 4 && sizeof(M) == 1, "int32 sums of a byte mask");
        qikkdb::check(qk_prefix_sum(reinterpret_cast<const int8_t*>(inMask), dataElementCount,
                                    reinterpret_cast<int32_t*>(prefixSumPointer), 1, nullptr));
    }

    template <typename M>
    static void PrefixSumExclusive(int32_t* prefixSumPointer, M* inMask, int32_t dataElementCount) {
        static_assert(sizeof(M) == 1, "masks are one byte per row");
        qikkdb::check(qk_prefix_sum(reinterpret_cast<const int8_t*>(inMask), dataElementCount,
                                    prefixSumPointer, 0, nullptr));
    }

    // Sum of a byte mask into a device scalar (SumKeep) or a host value (Sum).
    template <typename OUT, typename IN>
    static void SumKeep(OUT* outPointer, IN* inputBuffer, const int32_t dataElementCount) {
        OUT v{};
        Sum(v, inputBuffer, dataElementCount);
        GPUMemory::copyHostToDevice(outPointer, &v, 1);
    }

    template <typename OUT, typename IN>
    static void Sum(OUT& outValueHost, IN* inputBuffer, const int32_t dataElementCount) {
        static_assert(sizeof(IN) == 1, "SumKeep sums byte masks");
        int64_t s = 0;
        qikkdb::check(qk_mask_sum(reinterpret_cast<const int8_t*>(inputBuffer), dataElementCount, &s, nullptr));
        outValueHost = static_cast<OUT>(s);
    }

    static cuda_ptr<nullmask_t> CompressNullMask(nullarray_t* inputNullArray, int32_t dataElementCount) {
        uint64_t* out = nullptr;
        qikkdb::check(qk_compress_null_mask(inputNullArray, dataElementCount, &out, nullptr));
        return cuda_ptr<nullmask_t>(out);
    }
};
