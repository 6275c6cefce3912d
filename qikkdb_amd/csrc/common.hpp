// common.hpp - shared host/device helpers of the qikkdb_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/qikkdb_amd.h"

namespace qk {

// ---- error reporting -------------------------------------------------------------------
// Records `code` and a formatted message for qk_last_error_message() and returns `code`.
int fail(int code, const char* fmt, ...);

#define QK_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t qk_e_ = (call);                                                           \
        if (qk_e_ != hipSuccess)                                                             \
            return ::qk::fail(QK_RUNTIME_ERROR, "%s: %s (%s:%d)", #call,                     \
                              hipGetErrorString(qk_e_), __FILE__, __LINE__);                 \
    } while (0)

#define QK_TRY(expr)                                                                         \
    do {                                                                                     \
        int qk_s_ = (expr);                                                                  \
        if (qk_s_ != QK_SUCCESS) return qk_s_;                                               \
    } while (0)

// ---- device properties (of the device bound to the calling thread) ----------------------
struct DeviceInfo {
    int device = -1;
    int cus = 0;                 // compute units
    size_t lds_per_block = 0;    // bytes of LDS one workgroup may allocate
    size_t lds_per_cu = 0;
    int max_threads = 1024;
};
const DeviceInfo& device_info();
int ensure_init();

// ---- pooled allocation + per-stream scratch -------------------------------------------
int pool_alloc(void** p, size_t bytes);
int pool_free(void* p);
// Whether pool_alloc(bytes) would succeed under the pool's limit (or, without one, the device's
// free memory): a cached block of the size, or room under the limit.
bool pool_can_allocate(size_t bytes);

// Scratch owned by one stream: grows on demand, zero-filled when (re)allocated.  Holds the
// look-back status words, tickets and counters of the scan-based kernels; concurrent use
// from one stream is serialised by the stream itself.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint32_t epoch = 0;  // look-back epoch of the last launch on this scratch
};
int stream_scratch(hipStream_t s, size_t bytes, Scratch** out);

// Small pinned host buffer per thread for count read-backs.
int pinned_host(size_t bytes, void** out);

// Resolves and clears the current device's fixed scratch of qk_prefix_sum_noalloc (once per
// device; called by qk_init).
int prefix_noalloc_setup();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int type_size(int32_t t) {
    switch (t) {
    case QK_INT8: return 1;
    case QK_INT: return 4;
    case QK_FLOAT: return 4;
    case QK_LONG: return 8;
    case QK_DOUBLE: return 8;
    default: return 0;
    }
}
inline bool is_real_type(int32_t t) { return t == QK_FLOAT || t == QK_DOUBLE; }

// ---- device helpers ----------------------------------------------------------------------
__device__ __forceinline__ int null_bit(const uint64_t* __restrict__ m, int64_t i) {
    return (int)((m[i >> 6] >> (i & 63)) & 1ull);
}

// Order-preserving 64-bit encodings (unsigned order == value order) so that MIN/MAX of every
// value type runs on native unsigned 64-bit atomics.
__device__ __host__ __forceinline__ uint64_t enc_i64(int64_t x) {
    return (uint64_t)x ^ 0x8000000000000000ull;
}
__device__ __host__ __forceinline__ int64_t dec_i64(uint64_t u) {
    return (int64_t)(u ^ 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t enc_f64(double d) {
    uint64_t b = (uint64_t)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dec_f64(uint64_t u) {
    uint64_t b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double((long long)b);
}

// 64-bit key hash: a multiply by 2^64 / golden ratio (Fibonacci hashing: the high bits
// depend on every bit of the key), then the high half folded into the low half, so the low
// bits do too.  Tables take their sub-table from the high bits and the probe start from the
// low bits.
__device__ __host__ __forceinline__ uint64_t hash64(uint64_t k) {
    k *= 0x9e3779b97f4a7c15ull;
    return k ^ (k >> 32);
}

// Global-memory views of pointers that kernels load from memory (table views, argument
// structs).  The compiler cannot tell where such a pointer points and accesses it with flat
// instructions, which count against both the vector-memory and the LDS wait counters: a wait
// for any earlier load then drains every outstanding load.  Device allocations are global
// memory, so accesses through these views compile to global instructions instead.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_g(T* p) {
    return (gptr<T>)p;
}
// Relaxed, device-scope atomics on global memory (what atomicCAS / atomicAdd / atomicMin /
// atomicMax do on a generic pointer); each returns the old value.
__device__ __forceinline__ unsigned long long g_atomic_cas(unsigned long long* p, unsigned long long cmp,
                                                           unsigned long long val) {
    __hip_atomic_compare_exchange_strong(to_g(p), &cmp, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return cmp;
}
template <typename T>
__device__ __forceinline__ T g_atomic_add(T* p, T v) {
    return __hip_atomic_fetch_add(to_g(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long g_atomic_min(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_min(to_g(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long g_atomic_max(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_max(to_g(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T g_load_volatile(const T* p) {
    return *(const volatile __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-wide inclusive prefix sum of a 32-bit value (wave64).
__device__ __forceinline__ int32_t wave_incl_scan(int32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

}  // namespace qk
