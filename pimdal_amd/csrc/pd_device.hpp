// pd_device.hpp — device-side building blocks shared by the gfx950 kernels:
// wave64 lane masks and scans, agent-scope status words for decoupled
// look-back, and the reference hash functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PD_WAVE 64

namespace pd {

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

// Lanes strictly below this one.
__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << lane()) - 1ull;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }

// Buffer resource over `bytes` bytes from `base` (raw, untyped, 32-bit
// offsets): loads past the range return 0 and stores past it are dropped by
// the hardware range check.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// ---- agent-scope atomics (visible across the 8 XCDs' L2s) ----
template <typename T>
__device__ __forceinline__ T ld_acquire(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_release(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_relaxed(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- wave64 inclusive scan of uint32 via DPP-friendly shuffles ----
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t l = lane();
#pragma unroll
    for (int d = 1; d < PD_WAVE; d <<= 1) {
        uint32_t t = __shfl_up(v, d, PD_WAVE);
        if (l >= (uint32_t)d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
    const uint32_t l = lane();
#pragma unroll
    for (int d = 1; d < PD_WAVE; d <<= 1) {
        uint64_t t = __shfl_up(v, d, PD_WAVE);
        if (l >= (uint32_t)d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = PD_WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, PD_WAVE);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int d = PD_WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, PD_WAVE);
    return v;
}

// ---- single-value decoupled look-back (one running total per tile) ----
// Status word: bits 63..62 flag (0 empty, 1 tile aggregate, 2 inclusive
// prefix), bits 61..0 value.  One wave scans back over 64*W predecessors
// per step.
constexpr uint64_t LB_FLAG_AGG = 1ull << 62;
constexpr uint64_t LB_FLAG_PRE = 2ull << 62;
constexpr uint64_t LB_VALUE = (1ull << 62) - 1;

// Called by every lane of ONE wave of the tile's block.  Publishes the tile
// aggregate, accumulates the exclusive prefix and publishes the inclusive
// prefix.  Returns the exclusive prefix (same value in all lanes).
//
// Each step reads a window of 64*W predecessors (lane l reads tiles
// base - l*W - j, j < W, all loads in flight together), so a tile far from
// the nearest published inclusive prefix sums 64*W aggregates per coherent
// round trip.
//
// Split form: lb_publish_agg as soon as the tile's total is known (tile 0
// publishes its inclusive prefix directly), lb_resolve later.
__device__ __forceinline__ void lb_publish_agg(uint64_t* status, uint64_t tile, uint64_t agg) {
    if (lane() == 0) st_relaxed(&status[tile], (tile == 0 ? LB_FLAG_PRE : LB_FLAG_AGG) | agg);
}

template <int W = 1>
__device__ __forceinline__ uint64_t lb_resolve(uint64_t* status, uint64_t tile, uint64_t agg);

template <int W = 1>
__device__ __forceinline__ uint64_t lookback_single(uint64_t* status, uint64_t tile, uint64_t agg) {
    lb_publish_agg(status, tile, agg);
    return lb_resolve<W>(status, tile, agg);
}

// Walks back from an already published tile aggregate; publishes the
// inclusive prefix and returns the exclusive one.
template <int W>
__device__ __forceinline__ uint64_t lb_resolve(uint64_t* status, uint64_t tile, uint64_t agg) {
    if (tile == 0) return 0;
    uint64_t excl = 0;
    int64_t base = (int64_t)tile - 1;  // window covers [base - 64*W + 1, base]
    while (true) {
        uint64_t s[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            const int64_t idx = base - (int64_t)lane() * W - j;
            s[j] = (idx >= 0) ? ld_relaxed(&status[idx]) : LB_FLAG_PRE;  // before tile 0: prefix 0
        }
        // Wait until every status in the window is non-empty (all are
        // predecessors, already running, so each publishes soon).
        while (true) {
            bool empty = false;
#pragma unroll
            for (int j = 0; j < W; j++) empty |= (s[j] >> 62) == 0;
            if (!ballot(empty)) break;
            if (empty) {
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int j = 0; j < W; j++) {
                    if ((s[j] >> 62) == 0) s[j] = ld_relaxed(&status[base - (int64_t)lane() * W - j]);
                }
            }
        }
        // nearest inclusive prefix: first lane holding one, first j in that lane
        int first_j = W;
#pragma unroll
        for (int j = W - 1; j >= 0; j--) {
            if ((s[j] >> 62) == 2) first_j = j;
        }
        const uint64_t pre_mask = ballot(first_j < W);
        uint64_t v = 0;
        if (pre_mask) {
            const uint32_t fl = __builtin_ctzll(pre_mask);
#pragma unroll
            for (int j = 0; j < W; j++) {
                if (lane() < fl || (lane() == fl && j <= first_j)) v += s[j] & LB_VALUE;
            }
            excl += wave_sum64(v);
            break;
        }
#pragma unroll
        for (int j = 0; j < W; j++) v += s[j] & LB_VALUE;
        excl += wave_sum64(v);
        base -= (int64_t)PD_WAVE * W;
    }
    if (lane() == 0) st_relaxed(&status[tile], LB_FLAG_PRE | (excl + agg));
    return excl;
}

// ---- hashing ----
// reference/db_operators/hash_join.cpp:20-25: low 32 bits of the
// byte-swapped 64-bit product with 11400714785074694791 = 0x9E3779B185EBCA87.
__device__ __forceinline__ uint32_t ref_hash(uint32_t x) {
    uint64_t p = (uint64_t)x * 11400714785074694791ull;
    return (uint32_t)__builtin_bswap64(p);
}

// SplitMix64 finaliser (device copy; the host definition is in pimdal_gen.h)
__device__ __forceinline__ uint64_t pd_mix64_dev(uint64_t x) {
    x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27; x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

// Bucket inside a partition's table: bits of the same product that the
// partition number does not use.
__device__ __forceinline__ uint32_t slot_hash(uint32_t x) {
    uint64_t p = (uint64_t)x * 11400714785074694791ull;
    return (uint32_t)(p >> 16) ^ (uint32_t)(p >> 40);
}

}  // namespace pd
