// NullValues.h - null bit masks, as in QueryEngine/GPUCore/NullValues.cuh:6-45 of the
// reference: one bit per row, LSB first, in 64-bit words (row i is bit i % 64 of word
// i / 64; bit set = NULL).
#pragma once

#include <cstddef>
#include <cstdint>

typedef uint64_t nullmask_t;
typedef uint8_t nullarray_t;

class NullValues {
public:
    static int32_t GetBitMaskIdx(int32_t idx) { return idx / 64; }
    static int32_t GetShiftMaskIdx(int32_t idx) { return idx % 64; }
    static size_t GetNullBitMaskSize(size_t size) { return (size + 63) / 64; }
    static size_t GetNullBitMaskSizeInBytes(size_t size) { return GetNullBitMaskSize(size) * sizeof(nullmask_t); }
    static void SetBitInBitMask(nullmask_t* bitMask, int32_t bitMaskIdx, int32_t shiftMaskIdx, nullmask_t newBit) {
        bitMask[bitMaskIdx] = (bitMask[bitMaskIdx] & ~(nullmask_t(1) << shiftMaskIdx)) | ((newBit & 1) << shiftMaskIdx);
    }
    static void SetBitInBitMask(nullmask_t* bitMask, int32_t index, nullmask_t newBit) {
        SetBitInBitMask(bitMask, GetBitMaskIdx(index), GetShiftMaskIdx(index), newBit);
    }
    static nullmask_t GetConcreteBitFromBitmask(const nullmask_t* bitMask, int32_t bitMaskIdx, int32_t shiftMaskIdx) {
        return (bitMask[bitMaskIdx] >> shiftMaskIdx) & 1;
    }
    static nullmask_t GetConcreteBitFromBitmask(const nullmask_t* bitMask, int32_t index) {
        return GetConcreteBitFromBitmask(bitMask, GetBitMaskIdx(index), GetShiftMaskIdx(index));
    }
};
