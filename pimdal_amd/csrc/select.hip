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
    const __amdgpu_buffer_rsrc_t rp = sel_rsrc(pb, valid * 4u);
    const uint32_t l4 = pd::lane() * 4;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint32_t r0 = g * 256 + l4;
        if constexpr (VEC) {
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

// Persistent blocks take tiles in counter order, one tile ahead.  Per tile:
// predicate + wave ranks, kept rows staged compacted in the wave's LDS slice,
// then the next tile's loads are issued (into the registers just freed)
// before the look-back, so HBM reads stay in flight across it; finally each
// wave writes its kept rows with coalesced stores.  Forward progress: the
// lowest unpublished tile is always some block's current tile (a block only
// moves to its prefetched tile after publishing the current one), and all of
// its predecessors are published.
template <typename K, bool VEC, int LBW>
__global__ __launch_bounds__(SEL_BLOCK, sizeof(K) == 8 ? 4 : 5) void k_select(const K* __restrict__ key, const uint32_t* __restrict__ pay,
                                                      uint64_t n, uint64_t ulo, uint64_t wm1,
                                                      uint32_t* __restrict__ out_pay, uint32_t* __restrict__ out_rid,
                                                      uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
                                                      uint64_t* __restrict__ d_count, uint64_t ntiles) {
    constexpr int WAVES = SEL_BLOCK / PD_WAVE;
    __shared__ uint32_t s_tile[2];
    __shared__ uint32_t s_wave_tot[WAVES];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_pay[WAVES][SEL_WAVE_ROWS];
    __shared__ uint16_t s_idx[WAVES][SEL_WAVE_ROWS];

    const uint32_t t = threadIdx.x, l = pd::lane(), l4 = l * 4;
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / PD_WAVE);
    // rows of this wave in tile `tl` that exist (0..SEL_WAVE_ROWS), wave-uniform
    auto wave_rows = [&](uint32_t tl) -> uint32_t {
        const uint64_t wb = (uint64_t)tl * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
        return wb >= n ? 0u : (uint32_t)min((uint64_t)SEL_WAVE_ROWS, n - wb);
    };
    auto load = [&](uint32_t tl, K (&k)[SEL_GROUPS][4], uint32_t (&p)[SEL_GROUPS][4]) {
        const uint64_t wb = (uint64_t)tl * SEL_TILE + (uint64_t)w * SEL_WAVE_ROWS;
        sel_load<K, VEC>(key + wb, pay + wb, wave_rows(tl), k, p);
    };
    if (t == 0) s_tile[0] = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile[0]);
    if (tile >= ntiles) return;
    K k[SEL_GROUPS][4];
    uint32_t p[SEL_GROUPS][4];
    load(tile, k, p);
    uint32_t buf = 0;

    while (true) {
        if (t == 0) s_tile[buf ^ 1] = atomicAdd(tile_ctr, 1u);
        const uint32_t valid = wave_rows(tile);
        uint32_t bits = 0;
#pragma unroll
        for (int g = 0; g < SEL_GROUPS; g++)
#pragma unroll
            for (int e = 0; e < 4; e++)
                bits |= (uint32_t)((g * 256 + l4 + e < valid) && sel_keep<K>(k[g][e], ulo, wm1)) << (g * 4 + e);

        // ---- wave ranks: per group inclusive scan of packed 16-bit counts ----
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
                    s_pay[w][pos] = p[g][e];
                    s_idx[w][pos] = (uint16_t)(g * 256 + l4 + e);
                    pos++;
                }
            }
            wave_total += (uint32_t)(tot >> (16 * g)) & 0xFFFF;
        }
        if (l == 0) s_wave_tot[w] = wave_total;
        __syncthreads();
        const uint32_t next = __builtin_amdgcn_readfirstlane(s_tile[buf ^ 1]);
        if (next < ntiles) load(next, k, p);

        uint32_t wave_off = 0, tile_total = 0;
#pragma unroll
        for (int i = 0; i < WAVES; i++) {
            const uint32_t v = s_wave_tot[i];
            wave_off += (i < (int)w) ? v : 0;
            tile_total += v;
        }

        // ---- tile offset: one wave does the look-back ----
        if (w == 0) {
            const uint64_t excl = pd::lookback_single<LBW>(status, tile, tile_total);
            if (l == 0) {
                s_excl = excl;
                if (tile == ntiles - 1) *d_count = excl + tile_total;
            }
        }
        __syncthreads();

        // ---- coalesced writes of the wave's kept rows, in row order ----
        const uint64_t ob = s_excl + wave_off;
        const uint64_t obase = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(ob >> 32)) << 32) |
                               __builtin_amdgcn_readfirstlane((uint32_t)ob);
        const uint32_t row0 = tile * SEL_TILE + w * SEL_WAVE_ROWS;   // n < 2^32
        const uint32_t wt = __builtin_amdgcn_readfirstlane(wave_total);
        if (out_pay) {
            uint32_t* op = out_pay + obase;
            for (uint32_t i = l; i < wt; i += PD_WAVE) op[i] = s_pay[w][i];
        }
        if (out_rid) {
            uint32_t* orr = out_rid + obase;
            for (uint32_t i = l; i < wt; i += PD_WAVE) orr[i] = row0 + s_idx[w][i];
        }
        // No barrier needed before the next tile: each wave restages only its
        // own LDS slice (after its own writes), s_wave_tot is rewritten after
        // every wave passed the second barrier, and s_excl after the first
        // barrier of the next tile.
        if (next >= ntiles) break;
        tile = next;
        buf ^= 1;
    }
}

void select_range(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                  int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s) {
    PD_REQUIRE(kt == PIMDAL_U32 || kt == PIMDAL_I32 || kt == PIMDAL_I64, PIMDAL_E_UNSUPPORTED,
               "select key must be uint32, int32 or int64");
    PD_REQUIRE(kt != PIMDAL_I64 || pay || !out_pay, PIMDAL_E_INVALID, "int64 keys need a 32-bit payload column");
    PD_REQUIRE(n < (1ull << 32), PIMDAL_E_INVALID, "select supports up to 2^32-1 rows (uint32 row ids)");
    PD_REQUIRE(key || n == 0, PIMDAL_E_INVALID, "key is NULL");
    PD_REQUIRE(!out_pay || pay, PIMDAL_E_INVALID, "payload is NULL but out_payload requested");
    // without a payload column the key words stand in (read, never written out)
    if (!pay) pay = reinterpret_cast<const uint32_t*>(key);
    // clamp [lo, hi) to the key domain and express it as [ulo, ulo + wm1] in
    // the unsigned order of the key type
    uint64_t ulo = 0, wm1 = 0;
    bool empty = hi <= lo;
    if (!empty && kt != PIMDAL_I64) {
        const int64_t dmin = kt == PIMDAL_U32 ? 0 : -(1ll << 31);
        const int64_t dmax = dmin + (1ll << 32);   // exclusive
        const int64_t l2 = std::max(lo, dmin), h2 = std::min(hi, dmax);
        empty = h2 <= l2;
        ulo = (uint64_t)(l2 - dmin);
        wm1 = (uint64_t)(h2 - l2 - 1);
    } else if (!empty) {
        ulo = (uint64_t)lo ^ (1ull << 63);
        wm1 = (uint64_t)hi - (uint64_t)lo - 1;
    }
    if (n == 0 || empty) {
        fill_u64(c, d_count, 0, 1, s);
        return;
    }
    const uint64_t ntiles = ceil_div(n, SEL_TILE);
    Scratch scratch(c, 8 * (ntiles + 1), s);
    uint64_t* status = scratch.as<uint64_t>();
    uint32_t* ctr = reinterpret_cast<uint32_t*>(status + ntiles);
    fill_u64(c, status, 0, ntiles + 1, s);
    const bool vec = ((uintptr_t)key % 16 == 0) && ((uintptr_t)pay % 16 == 0);
    dim3 grid((uint32_t)std::min<uint64_t>(ntiles, (uint64_t)c->cu_count * (uint64_t)c->tune_get("select_blocks_per_cu", 5))),
        block(SEL_BLOCK);
    const int64_t lbw = c->tune_get("select_lookback_width", 4);
    auto go = [&](auto kp) {
        using K = std::remove_const_t<std::remove_pointer_t<decltype(kp)>>;
#define PD_SEL_LAUNCH(V, W) launch(c, "select", k_select<K, V, W>, grid, block, 0, s, kp, pay, n, ulo, wm1, out_pay, out_rid, status, ctr, d_count, ntiles)
        if (vec) {
            if (lbw >= 4) PD_SEL_LAUNCH(true, 4);
            else if (lbw >= 2) PD_SEL_LAUNCH(true, 2);
            else PD_SEL_LAUNCH(true, 1);
        } else {
            PD_SEL_LAUNCH(false, 4);
        }
#undef PD_SEL_LAUNCH
    };
    if (kt == PIMDAL_U32) go(reinterpret_cast<const uint32_t*>(key));
    else if (kt == PIMDAL_I32) go(reinterpret_cast<const int32_t*>(key));
    else go(reinterpret_cast<const int64_t*>(key));
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

extern "C" int pimdal_select(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* payload, int64_t lo, int64_t hi,
                             pimdal_col** out_payload, pimdal_col** out_rowid) {
    PD_API_BEGIN
    PD_REQUIRE(c && key, PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(!payload || payload->length == key->length, PIMDAL_E_INVALID, "payload length differs from key length");
    PD_REQUIRE(!payload || type_width(payload->type) == 4, PIMDAL_E_UNSUPPORTED, "payload must be 4 bytes wide");
    PD_REQUIRE(out_payload || out_rowid, PIMDAL_E_INVALID, "no output requested");
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
