// GPUMemory.h - device memory of the C++ drop-in layer (QueryEngine/GPUCore/GPUMemory.cuh:
// 89-194 of the reference): element-count allocation from the library's caching pool, set,
// copies, free.  Buffers returned by GetResults / *Keep are released with GPUMemory::free.
#pragma once

#include <cstddef>
#include <cstdint>

#include "GPUError.h"

namespace GPUMemory {

// A device string column (GPUMemory.cuh:70-77): every string's bytes in allChars, and
// stringIndices[i] = the end offset of string i ("start indices shifted by one string").
struct GPUString {
    char* allChars = nullptr;
    int64_t* stringIndices = nullptr;
};

template <typename T>
void alloc(T** p_Block, size_t dataElementCount) {
    void* p = nullptr;
    qikkdb::check(qk_malloc(&p, dataElementCount * sizeof(T)));
    *p_Block = static_cast<T*>(p);
}

template <typename T>
void memset(T* p_Block, int value, size_t dataElementCount) {
    qikkdb::check(qk_memset(p_Block, value, dataElementCount * sizeof(T), nullptr));
}

template <typename T>
void allocAndSet(T** p_Block, int value, size_t dataElementCount) {
    alloc(p_Block, dataElementCount);
    GPUMemory::memset(*p_Block, value, dataElementCount);
}

template <typename T>
void copyHostToDevice(T* p_BlockDevice, const T* p_BlockHost, size_t dataElementCount) {
    qikkdb::check(qk_memcpy_htod(p_BlockDevice, p_BlockHost, dataElementCount * sizeof(T), nullptr));
}

template <typename T>
void copyDeviceToHost(T* p_BlockHost, const T* p_BlockDevice, size_t dataElementCount) {
    qikkdb::check(qk_memcpy_dtoh(p_BlockHost, p_BlockDevice, dataElementCount * sizeof(T), nullptr));
}

inline void free(void* p_block) {
    if (p_block) qikkdb::check(qk_free(p_block));
}

inline void free(GPUString stringCol) {
    free(stringCol.allChars);
    free(stringCol.stringIndices);
}

}  // namespace GPUMemory
