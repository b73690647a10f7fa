// select.hip — order-preserving selection in one pass over HBM.
//
// Reference semantics: reference/db_operators/select.cpp:38-112 (count per
// block, prefix over blocks, compacting write) and pimdal/select/sel.c:47-128
// (per-tasklet compaction plus an ordered handshake prefix).  On gfx950 the
// two passes collapse into one: each 4096-row tile evaluates the predicate
// with 16-byte loads, ranks kept rows with wave64 ballots and popcounts,
// takes its output offset from a decoupled look-back over tile totals
// (tile ids come from an atomic counter, so every predecessor tile is
// already resident and the wait always ends), and writes the kept (payload,
// row id) pairs from registers into their contiguous output range.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int SEL_BLOCK = 256;
constexpr int SEL_GROUPS = 4;                              // uint4 groups per lane
constexpr int SEL_WAVE_ROWS = PD_WAVE * 4 * SEL_GROUPS;    // 1024
constexpr int SEL_TILE = SEL_BLOCK / PD_WAVE * SEL_WAVE_ROWS;  // 4096

// Loads one wave's share of a tile (row = base + g*256 + lane*4 + e) through
// buffer loads whose descriptor covers exactly the `valid` rows that exist:
// the hardware range check returns 0 past the end, so the ragged last tile
// needs no separate path and every address is a 32-bit lane offset from a
// wave-uniform base.  Rows at or past `valid` are masked by the predicate.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sel_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename K, bool VEC>
__device__ __forceinline__ void sel_load(const K* kb, const uint32_t* pb, uint32_t valid,
                                         K (&k)[SEL_GROUPS][4], uint32_t (&p)[SEL_GROUPS][4]) {
    const __amdgpu_buffer_rsrc_t rk = sel_rsrc(kb, valid * (uint32_t)sizeof(K));
    // no payload column: a descriptor of no bytes (every load returns 0 and
    // touches no memory)
    const __amdgpu_buffer_rsrc_t rp = sel_rsrc(pb, pb ? valid * 4u : 0u);
    const uint32_t l4 = pd::lane() * 4;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint32_t r0 = g * 256 + l4;
        if constexpr (sizeof(K) == 1) {
            // one dword per lane holds its four keys; a wave whose rows end
            // inside a dword loads bytes (the range check would return 0 for
            // the whole dword)
            if (VEC && valid == (uint32_t)SEL_WAVE_ROWS) {
                const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(rk, r0, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; e++) k[g][e] = (K)(v >> (8 * e));
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) k[g][e] = (K)__builtin_amdgcn_raw_buffer_load_b8(rk, r0 + e, 0, 0);
            }
            if constexpr (VEC) {
                const uint4 q = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rp, r0 * 4, 0, 0));
                p[g][0] = q.x; p[g][1] = q.y; p[g][2] = q.z; p[g][3] = q.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) p[g][e] = __builtin_amdgcn_raw_buffer_load_b32(rp, (r0 + e) * 4, 0, 0);
            }
        } else if constexpr (VEC) {
            if constexpr (sizeof(K) == 8) {
                const uint4 a = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 8, 0, 0));
                const uint4 b = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 8 + 16, 0, 0));
                k[g][0] = (K)(((uint64_t)a.y << 32) | a.x); k[g][1] = (K)(((uint64_t)a.w << 32) | a.z);
                k[g][2] = (K)(((uint64_t)b.y << 32) | b.x); k[g][3] = (K)(((uint64_t)b.w << 32) | b.z);
            } else {
                const uint4 v = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 4, 0, 0));
                k[g][0] = (K)v.x; k[g][1] = (K)v.y; k[g][2] = (K)v.z; k[g][3] = (K)v.w;
            }
            const uint4 q = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rp, r0 * 4, 0, 0));
            p[g][0] = q.x; p[g][1] = q.y; p[g][2] = q.z; p[g][3] = q.w;
        } else {
            // 4-byte aligned columns: one dword load per word
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if constexpr (sizeof(K) == 8) {
                    const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 8, 0, 0);
                    const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 8 + 4, 0, 0);
                    k[g][e] = (K)(((uint64_t)hi << 32) | lo);
                } else {
                    k[g][e] = (K)__builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 4, 0, 0);
                }
                p[g][e] = __builtin_amdgcn_raw_buffer_load_b32(rp, (r0 + e) * 4, 0, 0);
            }
        }
    }
}

// Predicate lo <= key < hi in the unsigned order of the key type:
// u = key bits with the sign bit flipped for signed keys, and a row is kept
// iff u - ulo <= wm1 (unsigned, wrapping), with [ulo, ulo + wm1] the range
// clamped to the key domain on the host.
template <typename K>
__device__ __forceinline__ bool sel_keep(K k, uint64_t ulo, uint64_t wm1) {
    if constexpr (sizeof(K) == 8) {
        const uint64_t u = (uint64_t)k ^ (1ull << 63);
        return u - ulo <= wm1;
    } else {
        const uint32_t u = (uint32_t)k ^ (std::is_signed<K>::value ? 0x80000000u : 0u);
        return (uint32_t)(u - (uint32_t)ulo) <= (uint32_t)wm1;
    }
}

// Predicate bits of one wave's 16 rows per lane, ranks within the wave
// (packed 16-bit group counts, one 64-bit wave scan), kept rows staged
// compacted in the wave's LDS slice.  Returns the wave's kept-row count.
template <typename K>
__device__ __forceinline__ uint32_t sel_count_stage(const K (&k)[SEL_GROUPS][4], const uint32_t (&p)[SEL_GROUPS][4],
                                                    uint32_t valid, uint64_t ulo, uint64_t wm1,
                                                    uint32_t* __restrict__ s_pay, uint16_t* __restrict__ s_idx) {
    const uint32_t l4 = pd::lane() * 4;
    uint32_t bits = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++)
#pragma unroll
        for (int e = 0; e < 4; e++)
            bits |= (uint32_t)((g * 256 + l4 + e < valid) && sel_keep<K>(k[g][e], ulo, wm1)) << (g * 4 + e);
    uint64_t packed = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) packed |= (uint64_t)__popc((bits >> (g * 4)) & 0xF) << (16 * g);
    const uint64_t incl = pd::wave_incl_scan64(packed);
    const uint64_t tot = __shfl(incl, PD_WAVE - 1, PD_WAVE);
    uint32_t wave_total = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint32_t gb = (bits >> (g * 4)) & 0xF;
        uint32_t pos = wave_total + ((uint32_t)(incl >> (16 * g)) & 0xFFFF) - __popc(gb);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (gb & (1u << e)) {
                s_pay[pos] = p[g][e];
                s_idx[pos] = (uint16_t)(g * 256 + l4 + e);
                pos++;
            }
        }
        wave_total += (uint32_t)(tot >> (16 * g)) & 0xFFFF;
    }
    return wave_total;
}

// Coalesced write of a wave's staged rows to [obase, obase + wt).
__device__ __forceinline__ void sel_write_staged(const uint32_t* __restrict__ s_pay, const uint16_t* __restrict__ s_idx,
                                                 uint32_t wt, uint64_t obase, uint32_t row0,
                                                 uint32_t* __restrict__ out_pay, uint32_t* __restrict__ out_rid) {
    const uint32_t l = pd::lane();
    if (out_pay) {
        uint32_t* op = out_pay + obase;
        for (uint32_t i = l; i < wt; i += PD_WAVE) op[i] = s_pay[i];
    }
    if (out_rid) {
        uint32_t* orr = out_rid + obase;
        for (uint32_t i = l; i < wt; i += PD_WAVE) orr[i] = row0 + s_idx[i];
    }
}

__device__ __forceinline__ uint64_t sel_uniform64(uint64_t v) {
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)v);
}

// Single pass.  Persistent blocks take tiles in counter order.  A tile's
// aggregate is published as soon as the tile is counted; its look-back is
// resolved one step later, while the block's next tile is loading:
//   take B -> load B -> resolve A (look-back) -> write A -> count B,
//   publish B's aggregate -> A = B.
// The current tile waits in LDS (kept rows compacted per wave) while the
// next one is in registers.  Forward progress: tile ids are taken in
// counter order, and a block publishes the aggregate of every tile it holds
// before it waits on any later look-back, so every predecessor of a tile in
// look-back has published or is about to.
template <typename K, bool VEC, int LBW, int BLOCK>
__global__ __launch_bounds__(BLOCK, (sizeof(K) == 8 || BLOCK == 1024) ? 4 : 5) void k_select(
    const K* __restrict__ key, const uint32_t* __restrict__ pay, uint64_t n, uint64_t ulo, uint64_t wm1,
    uint32_t* __restrict__ out_pay, uint32_t* __restrict__ out_rid, uint64_t* __restrict__ status,
    uint32_t* __restrict__ tile_ctr, uint64_t* __restrict__ d_count, uint64_t ntiles) {
    constexpr int WAVES = BLOCK / PD_WAVE;
    constexpr int TILE = WAVES * SEL_WAVE_ROWS;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_tot[WAVES];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_pay[WAVES][SEL_WAVE_ROWS];
    __shared__ uint16_t s_idx[WAVES][SEL_WAVE_ROWS];

    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / PD_WAVE);
    auto wave_rows = [&](uint32_t tl) -> uint32_t {
        const uint64_t wb = (uint64_t)tl * TILE + (uint64_t)w * SEL_WAVE_ROWS;
        return wb >= n ? 0u : (uint32_t)min((uint64_t)SEL_WAVE_ROWS, n - wb);
    };
    auto load = [&](uint32_t tl, K (&k)[SEL_GROUPS][4], uint32_t (&p)[SEL_GROUPS][4]) {
        const uint64_t wb = (uint64_t)tl * TILE + (uint64_t)w * SEL_WAVE_ROWS;
        sel_load<K, VEC>(key + wb, pay ? pay + wb : nullptr, wave_rows(tl), k, p);
    };
    K k[SEL_GROUPS][4];
    uint32_t p[SEL_GROUPS][4];

    // ---- prologue: first tile counted, staged and published ----
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
    if (tile >= ntiles) return;
    load(tile, k, p);
    uint32_t wt = sel_count_stage<K>(k, p, wave_rows(tile), ulo, wm1, s_pay[w], s_idx[w]);
    if (l == 0) s_wave_tot[w] = wt;
    __syncthreads();
    uint32_t woff = 0, ttot = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) {
        const uint32_t v = s_wave_tot[i];
        woff += (i < (int)w) ? v : 0;
        ttot += v;
    }
    if (w == 0) pd::lb_publish_agg(status, tile, ttot);

    while (true) {
        __syncthreads();   // every wave has read s_tile and s_wave_tot
        if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
        __syncthreads();
        const uint32_t next = __builtin_amdgcn_readfirstlane(s_tile);
        if (next < ntiles) load(next, k, p);

        if (w == 0) {
            const uint64_t excl = pd::lb_resolve<LBW>(status, tile, ttot);
            if (l == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) *d_count = excl + ttot;
            }
        }
        __syncthreads();
        sel_write_staged(s_pay[w], s_idx[w], wt, sel_uniform64(s_excl + woff), tile * TILE + w * SEL_WAVE_ROWS,
                         out_pay, out_rid);
        if (next >= ntiles) break;

        // the wave's own LDS slice is free again (its writes were issued first)
        wt = sel_count_stage<K>(k, p, wave_rows(next), ulo, wm1, s_pay[w], s_idx[w]);
        if (l == 0) s_wave_tot[w] = wt;
        __syncthreads();
        woff = 0;
        ttot = 0;
#pragma unroll
        for (int i = 0; i < WAVES; i++) {
            const uint32_t v = s_wave_tot[i];
            woff += (i < (int)w) ? v : 0;
            ttot += v;
        }
        tile = next;
        if (w == 0) pd::lb_publish_agg(status, tile, ttot);
    }
}

// Two passes (select_mode 1): per-block counts over contiguous row chunks,
// a scan of the counts, then a write pass over the same chunks.
constexpr int SEL2_CHUNK_TILES = 8;   // 4096-row tiles per block chunk

template <typename K, bool VEC>
__global__ __launch_bounds__(SEL_BLOCK) void k_select_count(const K* __restrict__ key, uint64_t n, uint64_t ulo,
                                                            uint64_t wm1, uint32_t* __restrict__ counts) {
    constexpr int WAVES = SEL_BLOCK / PD_WAVE;
    __shared__ uint32_t s_w[WAVES];
    const uint32_t t = threadIdx.x, l4 = pd::lane() * 4;
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / PD_WAVE);
    uint32_t cnt = 0;
    for (uint32_t j = 0; j < SEL2_CHUNK_TILES; j++) {
        const uint64_t wb = ((uint64_t)blockIdx.x * SEL2_CHUNK_TILES + j) * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
        const uint32_t valid = wb >= n ? 0u : (uint32_t)min((uint64_t)SEL_WAVE_ROWS, n - wb);
        if (!valid) break;
        const __amdgpu_buffer_rsrc_t rk = sel_rsrc(key + wb, valid * (uint32_t)sizeof(K));
#pragma unroll
        for (int g = 0; g < SEL_GROUPS; g++) {
            const uint32_t r0 = g * 256 + l4;
            K kk[4];
            if constexpr (sizeof(K) == 1) {
                if (VEC && valid == (uint32_t)SEL_WAVE_ROWS) {
                    const uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(rk, r0, 0, 0);
#pragma unroll
                    for (int e = 0; e < 4; e++) kk[e] = (K)(v >> (8 * e));
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++) kk[e] = (K)__builtin_amdgcn_raw_buffer_load_b8(rk, r0 + e, 0, 0);
                }
            } else if constexpr (VEC && sizeof(K) == 4) {
                const uint4 v = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 4, 0, 0));
                kk[0] = (K)v.x; kk[1] = (K)v.y; kk[2] = (K)v.z; kk[3] = (K)v.w;
            } else if constexpr (VEC) {
                const uint4 a = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 8, 0, 0));
                const uint4 b = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rk, r0 * 8 + 16, 0, 0));
                kk[0] = (K)(((uint64_t)a.y << 32) | a.x); kk[1] = (K)(((uint64_t)a.w << 32) | a.z);
                kk[2] = (K)(((uint64_t)b.y << 32) | b.x); kk[3] = (K)(((uint64_t)b.w << 32) | b.z);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if constexpr (sizeof(K) == 8) {
                        const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 8, 0, 0);
                        const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 8 + 4, 0, 0);
                        kk[e] = (K)(((uint64_t)hi << 32) | lo);
                    } else {
                        kk[e] = (K)__builtin_amdgcn_raw_buffer_load_b32(rk, (r0 + e) * 4, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 4; e++) cnt += (r0 + e < valid) && sel_keep<K>(kk[e], ulo, wm1);
        }
    }
    cnt = pd::wave_sum(cnt);
    if (pd::lane() == 0) s_w[w] = cnt;
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int i = 0; i < WAVES; i++) s += s_w[i];
        counts[blockIdx.x] = s;
    }
}

template <typename K, bool VEC>
__global__ __launch_bounds__(SEL_BLOCK, sizeof(K) == 8 ? 4 : 5) void k_select_write(
    const K* __restrict__ key, const uint32_t* __restrict__ pay, uint64_t n, uint64_t ulo, uint64_t wm1,
    const uint32_t* __restrict__ offs, uint32_t* __restrict__ out_pay, uint32_t* __restrict__ out_rid) {
    constexpr int WAVES = SEL_BLOCK / PD_WAVE;
    __shared__ uint32_t s_wave_tot[WAVES];
    __shared__ uint32_t s_pay[WAVES][SEL_WAVE_ROWS];
    __shared__ uint16_t s_idx[WAVES][SEL_WAVE_ROWS];
    const uint32_t t = threadIdx.x;
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t tile0 = (uint64_t)blockIdx.x * SEL2_CHUNK_TILES;
    uint64_t run = offs[blockIdx.x];
    auto wave_rows = [&](uint64_t tl) -> uint32_t {
        const uint64_t wb = tl * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
        return wb >= n ? 0u : (uint32_t)min((uint64_t)SEL_WAVE_ROWS, n - wb);
    };
    K k[SEL_GROUPS][4];
    uint32_t p[SEL_GROUPS][4];
    if (tile0 * SEL_TILE >= n) return;
    {
        const uint64_t wb = tile0 * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
        sel_load<K, VEC>(key + wb, pay ? pay + wb : nullptr, wave_rows(tile0), k, p);
    }
    for (uint32_t j = 0; j < SEL2_CHUNK_TILES; j++) {
        const uint64_t tl = tile0 + j;
        if (tl * SEL_TILE >= n) break;
        const uint32_t wt = sel_count_stage<K>(k, p, wave_rows(tl), ulo, wm1, s_pay[w], s_idx[w]);
        if (pd::lane() == 0) s_wave_tot[w] = wt;
        __syncthreads();
        const uint64_t nt = tl + 1;
        if (j + 1 < SEL2_CHUNK_TILES && nt * SEL_TILE < n) {
            const uint64_t wb = nt * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
            sel_load<K, VEC>(key + wb, pay ? pay + wb : nullptr, wave_rows(nt), k, p);
        }
        uint32_t woff = 0, ttot = 0;
#pragma unroll
        for (int i = 0; i < WAVES; i++) {
            const uint32_t v = s_wave_tot[i];
            woff += (i < (int)w) ? v : 0;
            ttot += v;
        }
        sel_write_staged(s_pay[w], s_idx[w], wt, sel_uniform64(run + woff), (uint32_t)(tl * SEL_TILE) + w * SEL_WAVE_ROWS,
                         out_pay, out_rid);
        run += ttot;
        __syncthreads();   // s_wave_tot is rewritten by the next tile
    }
}

// Host form of the predicate: [lo, hi) clamped to the key domain and
// expressed as [ulo, ulo + wm1] in the unsigned order of the key type.
// Returns false when no key value can satisfy it.
static bool sel_bounds(pimdal_type kt, int64_t lo, int64_t hi, uint64_t& ulo, uint64_t& wm1) {
    ulo = wm1 = 0;
    if (hi <= lo) return false;
    if (kt != PIMDAL_I64) {
        const int64_t dmin = (kt == PIMDAL_U32 || kt == PIMDAL_U8) ? 0 : -(1ll << 31);
        const int64_t dmax = dmin + (kt == PIMDAL_U8 ? 256 : (1ll << 32));   // exclusive
        const int64_t l2 = std::max(lo, dmin), h2 = std::min(hi, dmax);
        if (h2 <= l2) return false;
        ulo = (uint64_t)(l2 - dmin);
        wm1 = (uint64_t)(h2 - l2 - 1);
    } else {
        ulo = (uint64_t)lo ^ (1ull << 63);
        wm1 = (uint64_t)hi - (uint64_t)lo - 1;
    }
    return true;
}

static void sel_check_args(pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, const uint32_t* out_pay) {
    PD_REQUIRE(kt == PIMDAL_U32 || kt == PIMDAL_I32 || kt == PIMDAL_I64 || kt == PIMDAL_U8, PIMDAL_E_UNSUPPORTED,
               "select key must be uint8, uint32, int32 or int64");
    PD_REQUIRE(kt != PIMDAL_I64 || pay || !out_pay, PIMDAL_E_INVALID, "int64 keys need a 32-bit payload column");

    PD_REQUIRE(n < (1ull << 32), PIMDAL_E_INVALID, "select supports up to 2^32-1 rows (uint32 row ids)");
    PD_REQUIRE(key || n == 0, PIMDAL_E_INVALID, "key is NULL");
    PD_REQUIRE(!out_pay || pay, PIMDAL_E_INVALID, "payload is NULL but out_payload requested");
}

// One-pass block size for these columns: unaligned columns take the
// 256-thread kernel with dword loads.
static int sel_block(pimdal_ctx* c, bool vec) {
    const int64_t bsel = c->tune_get("select_block", 1024);
    return !vec ? 256 : bsel >= 1024 ? 1024 : bsel >= 512 ? 512 : 256;
}

static uint64_t sel_tiles(uint64_t n, int blk) { return ceil_div(n, (uint64_t)(blk / PD_WAVE) * SEL_WAVE_ROWS); }

__global__ void k_sel_clear(uint64_t* __restrict__ p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// The one-pass select over `status` (ntiles + 1 words: a look-back status
// per tile, then the tile counter), which it clears first.  PROFILED: the
// launches go through the context's hipEvent timing; otherwise they are
// plain launches, which create no events.
template <bool PROFILED>
static void select_one_pass(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n,
                            uint64_t ulo, uint64_t wm1, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count,
                            uint64_t* status, int BLK, bool vec, hipStream_t s) {
    const uint64_t ntl = sel_tiles(n, BLK);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(status + ntl);
    auto go_launch = [&](const char* name, auto kernel, dim3 grid, dim3 block, auto... args) {
        if constexpr (PROFILED) {
            launch(c, name, kernel, grid, block, 0, s, args...);
        } else {
            hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
            PD_HIP(hipGetLastError());
        }
    };
    go_launch("fill_u64", k_sel_clear, dim3((uint32_t)std::min<uint64_t>(ceil_div(ntl + 1, 256), 4096)), dim3(256),
              status, ntl + 1);
    const int64_t per_cu = c->tune_get("select_blocks_per_cu", BLK == 256 ? 5 : BLK == 512 ? 3 : 1);
    dim3 grid((uint32_t)std::min<uint64_t>(ntl, (uint64_t)c->cu_count * (uint64_t)per_cu));
    const int64_t lbw = c->tune_get("select_lookback_width", 1);
    auto go = [&](auto kp) {
        using K = std::remove_const_t<std::remove_pointer_t<decltype(kp)>>;
#define PD_SEL_LAUNCH(V, W, B) \
    go_launch("select", k_select<K, V, W, B>, grid, dim3(B), kp, pay, n, ulo, wm1, out_pay, out_rid, status, ctr, d_count, ntl)
        if (!vec) PD_SEL_LAUNCH(false, 1, 256);
        else if (BLK == 1024) PD_SEL_LAUNCH(true, 1, 1024);
        else if (BLK == 512) PD_SEL_LAUNCH(true, 1, 512);
        else if (lbw >= 4) PD_SEL_LAUNCH(true, 4, 256);
        else PD_SEL_LAUNCH(true, 1, 256);
#undef PD_SEL_LAUNCH
    };
    if (kt == PIMDAL_U32) go(reinterpret_cast<const uint32_t*>(key));
    else if (kt == PIMDAL_I32) go(reinterpret_cast<const int32_t*>(key));
    else if (kt == PIMDAL_U8) go(reinterpret_cast<const uint8_t*>(key));
    else go(reinterpret_cast<const int64_t*>(key));
}

void select_range(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                  int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s) {
    sel_check_args(kt, key, pay, n, out_pay);
    // without a payload column the key words stand in (read, never written
    // out); uint8 keys, a quarter of a payload's width, read none instead
    if (!pay && kt != PIMDAL_U8) pay = reinterpret_cast<const uint32_t*>(key);
    uint64_t ulo = 0, wm1 = 0;
    const bool empty = !sel_bounds(kt, lo, hi, ulo, wm1);
    if (n == 0 || empty) {
        fill_u64(c, d_count, 0, 1, s);
        return;
    }
    const uint64_t ntiles = ceil_div(n, SEL_TILE);
    const bool vec = ((uintptr_t)key % 16 == 0) && ((uintptr_t)pay % 16 == 0);
    if (c->tune_get("select_mode", 0) == 1) {
        // two passes: counts per block chunk, scan, write
        const uint64_t nb = ceil_div(ntiles, (uint64_t)SEL2_CHUNK_TILES);
        Scratch cnt(c, 4 * (nb + 1), s);
        auto go2 = [&](auto kp) {
            using K = std::remove_const_t<std::remove_pointer_t<decltype(kp)>>;
            if (vec) launch(c, "select_count", k_select_count<K, true>, dim3((uint32_t)nb), dim3(SEL_BLOCK), 0, s, kp, n, ulo, wm1, cnt.as<uint32_t>());
            else     launch(c, "select_count", k_select_count<K, false>, dim3((uint32_t)nb), dim3(SEL_BLOCK), 0, s, kp, n, ulo, wm1, cnt.as<uint32_t>());
            exclusive_scan_u32(c, cnt.as<uint32_t>(), cnt.as<uint32_t>(), nb, d_count, s);
            if (vec) launch(c, "select_write", k_select_write<K, true>, dim3((uint32_t)nb), dim3(SEL_BLOCK), 0, s, kp, pay, n, ulo, wm1, (const uint32_t*)cnt.as<uint32_t>(), out_pay, out_rid);
            else     launch(c, "select_write", k_select_write<K, false>, dim3((uint32_t)nb), dim3(SEL_BLOCK), 0, s, kp, pay, n, ulo, wm1, (const uint32_t*)cnt.as<uint32_t>(), out_pay, out_rid);
        };
        if (kt == PIMDAL_U32) go2(reinterpret_cast<const uint32_t*>(key));
        else if (kt == PIMDAL_I32) go2(reinterpret_cast<const int32_t*>(key));
        else if (kt == PIMDAL_U8) go2(reinterpret_cast<const uint8_t*>(key));
        else go2(reinterpret_cast<const int64_t*>(key));
        return;
    }
    // one pass: tile = block size / 64 waves x 1024 rows
    const int BLK = sel_block(c, vec);
    Scratch scratch(c, 8 * (sel_tiles(n, BLK) + 1), s);
    select_one_pass<true>(c, kt, key, pay, n, ulo, wm1, out_pay, out_rid, d_count, scratch.as<uint64_t>(), BLK, vec, s);
}

// ---- no-allocation select ----
// The one-pass select's status words and tile counter live in `scratch`, a
// file-scope device array of SEL_NA_WORDS words (1 MiB + 256 B, held by
// every GPU the library runs on until the process ends), instead of
// memory from the context's pool.  That bounds a call to 2^29 rows:
// 2^17 tiles of the smallest (4096-row) tile shape plus the counter.
constexpr uint64_t SEL_NA_WORDS = (1ull << 17) + 32;
constexpr uint64_t SEL_NA_MAX_ROWS = 1ull << 29;
static __device__ __align__(256) uint64_t scratch[SEL_NA_WORDS];

// The array's device address on each device.  The name `scratch` in host
// code is only the runtime's handle for the symbol, never a device address.
constexpr int SEL_NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static uint64_t* g_na_ptr[SEL_NA_MAX_DEVICES] = {};

// Looks up `scratch`'s address on `device`, which must be the current
// device; once per device (later calls return at once).
void select_noalloc_init(int device) {
    PD_REQUIRE(device >= 0 && device < SEL_NA_MAX_DEVICES, PIMDAL_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return;
    void* p = nullptr;
    PD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PD_REQUIRE(p != nullptr, PIMDAL_E_HIP, "no device address for the select scratch array");
    g_na_ptr[device] = static_cast<uint64_t*>(p);
}

static uint64_t* select_noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_na_mu);
    return (device >= 0 && device < SEL_NA_MAX_DEVICES) ? g_na_ptr[device] : nullptr;
}

void select_range_noalloc(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                          int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s) {
    sel_check_args(kt, key, pay, n, out_pay);
    PD_REQUIRE(n <= SEL_NA_MAX_ROWS, PIMDAL_E_CAPACITY,
               "pimdal_select_dev_noalloc takes at most 2^29 rows (its static scratch holds 2^17 tile statuses)");
    uint64_t* status = select_noalloc_scratch(c->device);
    PD_REQUIRE(status != nullptr, PIMDAL_E_INTERNAL, "select scratch address not set up for this device");
    if (!pay && kt != PIMDAL_U8) pay = reinterpret_cast<const uint32_t*>(key);
    uint64_t ulo = 0, wm1 = 0;
    const bool empty = !sel_bounds(kt, lo, hi, ulo, wm1);
    if (n == 0 || empty) {
        hipLaunchKernelGGL(k_sel_clear, dim3(1), dim3(64), 0, s, d_count, (uint64_t)1);
        PD_HIP(hipGetLastError());
        return;
    }
    const bool vec = ((uintptr_t)key % 16 == 0) && ((uintptr_t)pay % 16 == 0);
    const int BLK = sel_block(c, vec);
    PD_REQUIRE(sel_tiles(n, BLK) + 1 <= SEL_NA_WORDS, PIMDAL_E_INTERNAL, "select scratch too small");
    select_one_pass<false>(c, kt, key, pay, n, ulo, wm1, out_pay, out_rid, d_count, status, BLK, vec, s);
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_select_dev(pimdal_ctx* c, pimdal_type key_type, const void* key, const uint32_t* payload,
                                 uint64_t n, int64_t lo, int64_t hi, uint32_t* out_payload, uint32_t* out_rowid,
                                 uint64_t* d_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && d_count, PIMDAL_E_INVALID, "NULL argument");
    select_range(c, key_type, key, payload, n, lo, hi, out_payload, out_rowid, d_count, c->stream);
    PD_API_END
}

extern "C" int pimdal_select_dev_noalloc(pimdal_ctx* c, pimdal_type key_type, const void* key,
                                         const uint32_t* payload, uint64_t n, int64_t lo, int64_t hi,
                                         uint32_t* out_payload, uint32_t* out_rowid, uint64_t* d_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && d_count, PIMDAL_E_INVALID, "NULL argument");
    select_range_noalloc(c, key_type, key, payload, n, lo, hi, out_payload, out_rowid, d_count, c->stream);
    PD_API_END
}

extern "C" int pimdal_select(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* payload, int64_t lo, int64_t hi,
                             pimdal_col** out_payload, pimdal_col** out_rowid) {
    PD_API_BEGIN
    PD_REQUIRE(c && key, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    PD_REQUIRE(!payload || payload->length == key->length, PIMDAL_E_INVALID, "payload length differs from key length");
    PD_REQUIRE(!payload || type_width(payload->type) == 4, PIMDAL_E_UNSUPPORTED, "payload must be 4 bytes wide");
    PD_REQUIRE(out_payload || out_rowid, PIMDAL_E_INVALID, "no output requested");
    if (key->validity || (payload && payload->validity)) {
        // a NULL key satisfies no predicate; kept payloads keep their validity
        select_cols_nullable(c, key, payload, lo, hi, out_payload, out_rowid);
        return PIMDAL_OK;
    }
    const uint64_t n = key->length;
    const pimdal_type pt = payload ? payload->type : key->type;
    pimdal_col* op = nullptr;
    pimdal_col* orow = nullptr;
    Scratch cnt(c, 8);
    try {
        if (out_payload) op = new pimdal_col{pt, 0, n, c->pool.alloc(std::max<uint64_t>(n, 1) * 4, c->stream), true};
        if (out_rowid) orow = new pimdal_col{PIMDAL_U32, 0, n, c->pool.alloc(std::max<uint64_t>(n, 1) * 4, c->stream), true};
        select_range(c, key->type, key->data, payload ? (const uint32_t*)payload->data : nullptr, n, lo, hi,
                     op ? (uint32_t*)op->data : nullptr, orow ? (uint32_t*)orow->data : nullptr,
                     cnt.as<uint64_t>(), c->stream);
        const uint64_t k = read_u64(c, cnt.get(), c->stream);
        if (op) op->length = k;
        if (orow) orow->length = k;
    } catch (...) {
        if (op) { c->pool.free(op->data, c->stream); delete op; }
        if (orow) { c->pool.free(orow->data, c->stream); delete orow; }
        throw;
    }
    if (out_payload) *out_payload = op;
    if (out_rowid) *out_rowid = orow;
    PD_API_END
}
