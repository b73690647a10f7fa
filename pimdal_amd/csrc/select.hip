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

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int SEL_BLOCK = 256;
constexpr int SEL_GROUPS = 4;                              // uint4 groups per lane
constexpr int SEL_WAVE_ROWS = PD_WAVE * 4 * SEL_GROUPS;    // 1024
constexpr int SEL_TILE = SEL_BLOCK / PD_WAVE * SEL_WAVE_ROWS;  // 4096

template <typename K, bool VEC>
__global__ __launch_bounds__(SEL_BLOCK) void k_select(const K* __restrict__ key, const uint32_t* __restrict__ pay,
                                                      uint64_t n, int64_t lo, int64_t hi,
                                                      uint32_t* __restrict__ out_pay, uint32_t* __restrict__ out_rid,
                                                      uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
                                                      uint64_t* __restrict__ d_count, uint64_t ntiles) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_tot[SEL_BLOCK / PD_WAVE];
    __shared__ uint64_t s_excl;

    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    // persistent: each block takes tiles in counter order until none are left
    // (a tile's predecessors were taken earlier, by running blocks)
    while (true) {
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    if (tile >= ntiles) break;
    const uint64_t tile_base = tile * SEL_TILE;
    const uint64_t wave_base = tile_base + (uint64_t)w * SEL_WAVE_ROWS;

    // ---- load keys and payloads together (row = wave_base + g*256 + l*4 + e) ----
    K k[SEL_GROUPS][4];
    uint32_t p[SEL_GROUPS][4];
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint64_t r0 = wave_base + (uint64_t)g * 256 + l * 4;
        if (VEC && r0 + 3 < n) {
            if constexpr (sizeof(K) == 8) {
                const uint4 a = *reinterpret_cast<const uint4*>(key + r0);
                const uint4 b = *reinterpret_cast<const uint4*>(key + r0 + 2);
                k[g][0] = (K)(((uint64_t)a.y << 32) | a.x); k[g][1] = (K)(((uint64_t)a.w << 32) | a.z);
                k[g][2] = (K)(((uint64_t)b.y << 32) | b.x); k[g][3] = (K)(((uint64_t)b.w << 32) | b.z);
            } else {
                const uint4 v = *reinterpret_cast<const uint4*>(key + r0);
                k[g][0] = (K)v.x; k[g][1] = (K)v.y; k[g][2] = (K)v.z; k[g][3] = (K)v.w;
            }
            const uint4 q = *reinterpret_cast<const uint4*>(pay + r0);
            p[g][0] = q.x; p[g][1] = q.y; p[g][2] = q.z; p[g][3] = q.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                k[g][e] = (r0 + e < n) ? key[r0 + e] : (K)0;
                p[g][e] = (r0 + e < n) ? pay[r0 + e] : 0u;
            }
        }
    }
    uint32_t bits = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint64_t r0 = wave_base + (uint64_t)g * 256 + l * 4;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const bool in = (r0 + e < n) && ((int64_t)k[g][e] >= lo) && ((int64_t)k[g][e] < hi);
            bits |= (uint32_t)in << (g * 4 + e);
        }
    }

    // ---- wave ranks: per group inclusive scan of packed 16-bit counts ----
    uint64_t packed = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) packed |= (uint64_t)__popc((bits >> (g * 4)) & 0xF) << (16 * g);
    const uint64_t incl = pd::wave_incl_scan64(packed);
    const uint64_t tot = __shfl(incl, PD_WAVE - 1, PD_WAVE);
    uint32_t group_off[SEL_GROUPS];
    uint32_t wave_total = 0;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint32_t gi = (uint32_t)(incl >> (16 * g)) & 0xFFFF;
        const uint32_t gt = (uint32_t)(tot >> (16 * g)) & 0xFFFF;
        const uint32_t mine = __popc((bits >> (g * 4)) & 0xF);
        group_off[g] = wave_total + gi - mine;  // exclusive, within the wave
        wave_total += gt;
    }
    if (l == 0) s_wave_tot[w] = wave_total;
    __syncthreads();
    uint32_t wave_off = 0, tile_total = 0;
#pragma unroll
    for (int i = 0; i < SEL_BLOCK / PD_WAVE; i++) {
        const uint32_t v = s_wave_tot[i];
        wave_off += (i < (int)w) ? v : 0;
        tile_total += v;
    }

    // ---- tile offset: one wave does the look-back ----
    if (w == 0) {
        const uint64_t excl = pd::lookback_single(status, tile, tile_total);
        if (l == 0) {
            s_excl = excl;
            if (tile == ntiles - 1) *d_count = excl + tile_total;
        }
    }
    __syncthreads();

    // ---- write kept rows straight from registers, in row order ----
    const uint64_t base = s_excl + wave_off;
#pragma unroll
    for (int g = 0; g < SEL_GROUPS; g++) {
        const uint32_t gb = (bits >> (g * 4)) & 0xF;
        uint64_t o = base + group_off[g];
        const uint64_t r0 = wave_base + (uint64_t)g * 256 + l * 4;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (gb & (1u << e)) {
                if (o < n) {
                    if (out_pay) out_pay[o] = p[g][e];
                    if (out_rid) out_rid[o] = (uint32_t)(r0 + e);
                }
                o++;
            }
        }
    }
    __syncthreads();  // shared tile id / totals are rewritten by the next iteration
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
    if (n == 0) {
        fill_u64(c, d_count, 0, 1, s);
        return;
    }
    const uint64_t ntiles = ceil_div(n, SEL_TILE);
    Scratch scratch(c, 8 * (ntiles + 1), s);
    uint64_t* status = scratch.as<uint64_t>();
    uint32_t* ctr = reinterpret_cast<uint32_t*>(status + ntiles);
    fill_u64(c, status, 0, ntiles + 1, s);
    const bool vec = ((uintptr_t)key % 16 == 0) && ((uintptr_t)pay % 16 == 0);
    dim3 grid((uint32_t)std::min<uint64_t>(ntiles, (uint64_t)c->cu_count * (uint64_t)c->tune_get("select_blocks_per_cu", 8))),
        block(SEL_BLOCK);
    if (kt == PIMDAL_U32) {
        auto* kp = reinterpret_cast<const uint32_t*>(key);
        if (vec) launch(c, "select", k_select<uint32_t, true>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
        else     launch(c, "select", k_select<uint32_t, false>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
    } else if (kt == PIMDAL_I32) {
        auto* kp = reinterpret_cast<const int32_t*>(key);
        if (vec) launch(c, "select", k_select<int32_t, true>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
        else     launch(c, "select", k_select<int32_t, false>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
    } else {
        auto* kp = reinterpret_cast<const int64_t*>(key);
        if (vec) launch(c, "select", k_select<int64_t, true>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
        else     launch(c, "select", k_select<int64_t, false>, grid, block, 0, s, kp, pay, n, lo, hi, out_pay, out_rid, status, ctr, d_count, ntiles);
    }
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
