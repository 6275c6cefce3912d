// lookback.hpp - single-pass tile prefix by decoupled look-back
// (Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016).
//
// Forward progress: tiles are numbered by an atomic ticket in the order workgroups start, and
// a tile only waits for tiles with smaller tickets, which have already started and do not
// wait for it.  Status words carry an epoch so that the array never needs clearing between
// launches: a word from an earlier launch reads as "not ready".
#pragma once

#include "common.hpp"

namespace qk {

// status word: [63:62] flag, [61:32] epoch (30 bits), [31:0] value (modulo 2^32)
constexpr uint64_t kFlagAggregate = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;

__device__ __forceinline__ uint64_t lb_pack(uint64_t flag, uint32_t epoch, uint32_t v) {
    return flag | ((uint64_t)(epoch & 0x3fffffffu) << 32) | (uint64_t)v;
}

// Scratch layout used by every look-back kernel: word 0 holds the ticket counter, status
// words start at word 8 (a separate 64-byte line).
constexpr size_t kLbHeaderWords = 8;
inline size_t lb_scratch_bytes(int64_t tiles) { return (size_t)(kLbHeaderWords + tiles + 1) * 8; }

// Thread 0 draws the tile number; the holder of the last ticket re-arms the counter for the
// next launch (every ticket of this launch has been drawn by then).
__device__ __forceinline__ int64_t lb_ticket(uint32_t* ticket, int64_t tiles) {
    uint32_t t = atomicAdd(ticket, 1u);
    if ((int64_t)t == tiles - 1)
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (int64_t)t;
}

// Called by all 64 lanes of ONE wave.  Publishes `aggregate` for `tile`, looks back for the
// exclusive prefix, publishes the inclusive prefix and returns the exclusive prefix on
// every lane.
__device__ __forceinline__ uint32_t lb_exclusive(uint64_t* status, uint32_t epoch, int64_t tile,
                                                 uint32_t aggregate) {
    const int lane = threadIdx.x & 63;
    if (tile == 0) {
        if (lane == 0) st_relaxed(&status[0], lb_pack(kFlagPrefix, epoch, aggregate));
        return 0;
    }
    if (lane == 0) st_relaxed(&status[tile], lb_pack(kFlagAggregate, epoch, aggregate));
    uint32_t excl = 0;
    int64_t base = tile - 1;
    const uint32_t ep = epoch & 0x3fffffffu;
    for (;;) {
        const int64_t idx = base - lane;
        uint64_t w = idx >= 0 ? ld_relaxed(&status[idx]) : lb_pack(kFlagPrefix, epoch, 0);
        const uint64_t flag = w & (3ull << 62);
        const bool valid = (((w >> 32) & 0x3fffffffu) == ep) && flag != 0;
        const uint64_t pmask = __ballot(valid && flag == kFlagPrefix);
        const uint64_t bad = __ballot(!valid);
        uint64_t need;  // lanes whose values are summed this round
        if (pmask) {
            const int fp = __builtin_ctzll(pmask);
            need = (fp == 63) ? ~0ull : ((2ull << fp) - 1);
        } else {
            need = ~0ull;
        }
        if (bad & need) {  // a predecessor has not published yet: read the window again
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint32_t v = ((need >> lane) & 1ull) ? (uint32_t)w : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        excl += v;
        if (pmask) break;
        base -= 64;
    }
    if (lane == 0) st_relaxed(&status[tile], lb_pack(kFlagPrefix, epoch, excl + aggregate));
    return excl;
}

}  // namespace qk
