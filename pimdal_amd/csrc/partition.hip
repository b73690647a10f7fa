// partition.hip — one-pass hash partition into up to 1024 partitions.
//
// Reference: reference/db_operators/hash_join.cpp:57-112 partitions both
// relations by hash % part_n in two passes (count, then write); the PIM
// join radix-partitions over several levels (pimdal/join/hash_join.c:154-235,
// part_kernel).  Here one pass moves every row once, so the join pays one
// partition pass per side instead of the two that 8-bit radix passes need
// for more than 256 partitions:
//  k_pw_count    per tile (8192 or 16384 rows), rows per partition (LDS
//                counters);
//  k_pw_tscan    per partition, the rows of earlier tiles (a scan over tiles
//                in groups of PW_GT, groups chained by decoupled look-back),
//                and the partition starts (= bounds);
//  k_pw_scatter  per tile: rows ranked within their partition by LDS atomic
//                counters (order inside a partition is unspecified), staged
//                in LDS by partition, written out as (key, payload) rows in
//                partition runs (16 rows, 128 bytes, per partition per tile
//                on average at 1024 partitions),
//                tiles dealt to the 8 XCDs in contiguous runs so neighbouring
//                tiles' runs meet in one L2 and leave it as whole lines.
// Partition of a row: ref_hash(key) & (F - 1), the numbering of the radix
// passes over hash digits (hash_partition).
#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr uint32_t PW_BLOCK = 1024;
constexpr uint32_t PW_BINS = 1024;
constexpr uint32_t PW_GT = 32;                      // tiles per scan group
constexpr uint32_t PW_FLAG_AGG = 1u << 30;
constexpr uint32_t PW_FLAG_PRE = 2u << 30;
constexpr uint32_t PW_VALUE = (1u << 30) - 1;

// Persistent: block b counts tiles b, b + gridDim.x, ... (TILE rows each,
// 512 threads x TILE / 512 rows).
template <uint32_t TILE>
__global__ __launch_bounds__(512) void k_pw_count(const uint32_t* __restrict__ kin, uint64_t n, uint32_t mask,
                                                  uint32_t* __restrict__ cnt) {
    constexpr uint32_t Q = TILE / 512;   // rows per thread
    __shared__ uint32_t h[PW_BINS];
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t ntiles = (n + TILE - 1) / TILE;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        for (uint32_t i = t; i < PW_BINS; i += 512) h[i] = 0;
        const uint64_t wave_base = tile * TILE + (uint64_t)w * (PD_WAVE * Q);
        const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
            (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * Q), n - wave_base) : 0));
        const __amdgpu_buffer_rsrc_t rk = buf_rsrc(kin + wave_base, wrows * 4u);
        uint32_t q[Q];
#pragma unroll
        for (uint32_t i = 0; i < Q; i++) q[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < Q; i++)
            if (l + i * PD_WAVE < wrows) atomicAdd(&h[ref_hash(q[i]) & mask], 1u);
        __syncthreads();
        for (uint32_t i = t; i < PW_BINS; i += 512) cnt[tile * PW_BINS + i] = h[i];
        __syncthreads();
    }
}

// In place: cnt[tile][d] becomes the rows of partition d in tiles before
// `tile`; start[d] (d <= F) = rows of partitions below d.  Thread t =
// partition t; a block takes the next group of PW_GT tiles from a counter.
__global__ __launch_bounds__(PW_BINS) void k_pw_tscan(uint32_t* __restrict__ cnt, uint32_t ntiles,
                                                      uint32_t* __restrict__ gstatus, uint32_t* __restrict__ ctr,
                                                      uint32_t* __restrict__ start, uint32_t F) {
    __shared__ uint32_t s_g, s_w[PW_BINS / PD_WAVE];
    const uint32_t t = threadIdx.x, l = lane(), w = t / PD_WAVE;
    if (t == 0) s_g = atomicAdd(ctr, 1u);
    __syncthreads();
    const uint32_t g = s_g;
    const uint32_t ngroups = (ntiles + PW_GT - 1) / PW_GT;
    const uint32_t t0 = g * PW_GT, t1 = min(ntiles, t0 + PW_GT);
    // the group's counts, all loads in flight together
    uint32_t cg[PW_GT], sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < PW_GT; i++) cg[i] = t0 + i < t1 ? cnt[(uint64_t)(t0 + i) * PW_BINS + t] : 0u;
#pragma unroll
    for (uint32_t i = 0; i < PW_GT; i++) sum += cg[i];
    st_relaxed(&gstatus[(uint64_t)g * PW_BINS + t], (g == 0 ? PW_FLAG_PRE : PW_FLAG_AGG) | sum);
    uint32_t excl = 0;
    if (g > 0) {
        int64_t j = (int64_t)g - 1;
        bool done = false;
        while (!done) {   // every earlier group is running (tickets), so each publishes
            uint32_t st[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                st[q] = (j - q >= 0) ? ld_relaxed(&gstatus[(uint64_t)(j - q) * PW_BINS + t]) : PW_FLAG_PRE;
            bool stall = false;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (done || stall) continue;
                if ((st[q] >> 30) == 0) {
                    stall = true;
                    continue;
                }
                excl += st[q] & PW_VALUE;
                if ((st[q] >> 30) == 2) done = true;
                else j--;
            }
            if (stall) __builtin_amdgcn_s_sleep(1);
        }
        st_relaxed(&gstatus[(uint64_t)g * PW_BINS + t], PW_FLAG_PRE | (excl + sum));
    }
    uint32_t run = excl;
#pragma unroll
    for (uint32_t i = 0; i < PW_GT; i++) {
        if (t0 + i < t1) cnt[(uint64_t)(t0 + i) * PW_BINS + t] = run;
        run += cg[i];
    }
    if (g == ngroups - 1) {
        // run = rows of partition t: exclusive scan over partitions
        const uint32_t inc = wave_incl_scan(run);
        if (l == PD_WAVE - 1) s_w[w] = inc;
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t i = 0; i < w; i++) add += s_w[i];
        if (t < F) start[t] = add + inc - run;
        if (t == PW_BINS - 1) start[F] = add + inc;   // every row
    }
}

// Block b takes tile (b % 8) * xt + b / 8 (xt = tiles per XCD): an XCD's
// blocks work on neighbouring tiles together, so partition runs of
// neighbouring tiles meet in its L2.  ITEMS rows per thread: 16 (16384-row
// tiles, 136 KB of LDS, one block per CU) or 8 (8192-row tiles, 72 KB, two
// blocks per CU, one loading while the other writes).
template <uint32_t ITEMS>
__global__ __launch_bounds__(PW_BLOCK) void k_pw_scatter(const uint32_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin, uint2* __restrict__ out,
                                                         uint64_t n, uint32_t mask,
                                                         const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ start, uint32_t xt) {
    constexpr uint32_t TILE = PW_BLOCK * ITEMS;
    __shared__ __align__(16) uint32_t s_k[TILE];
    __shared__ __align__(16) uint32_t s_v[TILE];
    __shared__ uint32_t s_c[PW_BINS];   // tile rows per partition, then their start in the tile
    __shared__ uint32_t s_g[PW_BINS];   // output position of the partition's run minus its tile start
    __shared__ uint32_t s_ws[PW_BLOCK / PD_WAVE];
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t tile = (uint64_t)(blockIdx.x % 8) * xt + blockIdx.x / 8;
    const uint64_t tile_base = tile * TILE;
    if (tile_base >= n) return;   // grid rounded up to a multiple of 8
    const uint32_t tile_rows = (uint32_t)min((uint64_t)TILE, n - tile_base);
    s_c[t] = 0;
    const uint32_t gpos = (t <= mask ? start[t] : 0u) + cnt[tile * PW_BINS + t];
    const uint64_t wave_base = tile_base + (uint64_t)w * (PD_WAVE * ITEMS);
    const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
        (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * ITEMS), n - wave_base) : 0));
    const __amdgpu_buffer_rsrc_t rk = buf_rsrc(kin + wave_base, wrows * 4u);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vin + wave_base, wrows * 4u);
    uint32_t key[ITEMS], val[ITEMS];
#pragma unroll
    for (uint32_t i = 0; i < ITEMS; i++) {
        key[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
        val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, (l + i * PD_WAVE) * 4u, 0, 0);
    }
    __syncthreads();
    // rank within the partition (LDS counter), partition in the high bits
    uint32_t rk_d[ITEMS];
#pragma unroll
    for (uint32_t i = 0; i < ITEMS; i++) {
        rk_d[i] = 0xFFFFFFFFu;
        if (l + i * PD_WAVE < wrows) {
            const uint32_t d = ref_hash(key[i]) & mask;
            rk_d[i] = atomicAdd(&s_c[d], 1u) | (d << 16);
        }
    }
    __syncthreads();
    // tile-local partition starts (exclusive scan over the 1024 counters)
    const uint32_t c = s_c[t];
    const uint32_t inc = wave_incl_scan(c);
    if (l == PD_WAVE - 1) s_ws[w] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t i = 0; i < w; i++) add += s_ws[i];
    const uint32_t st = add + inc - c;
    s_c[t] = st;
    s_g[t] = gpos - st;
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < ITEMS; i++) {
        if (rk_d[i] == 0xFFFFFFFFu) continue;
        const uint32_t j = s_c[rk_d[i] >> 16] + (rk_d[i] & 0xFFFFu);
        s_k[j] = key[i];
        s_v[j] = val[i];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < ITEMS; i++) {
        const uint32_t j = i * PW_BLOCK + t;
        if (j >= tile_rows) break;
        const uint32_t k = s_k[j];
        out[(uint64_t)s_g[ref_hash(k) & mask] + j] = make_uint2(k, s_v[j]);
    }
}

void partition_tile_scan(pimdal_ctx* c, uint32_t* cnt, uint32_t ntiles, uint32_t F, uint32_t* start, hipStream_t s) {
    PD_REQUIRE(F >= 1 && F <= PW_BINS, PIMDAL_E_INVALID, "tile scan: 1 to 1024 partitions");
    const uint32_t ngroups = (ntiles + PW_GT - 1) / PW_GT;
    Scratch gst(c, ((uint64_t)ngroups * PW_BINS + 1) * 4, s);
    fill_u32(c, gst.get(), 0, (uint64_t)ngroups * PW_BINS + 1, s);
    launch(c, "pw_tscan", k_pw_tscan, dim3(ngroups), dim3(PW_BINS), 0, s, cnt, ntiles, gst.as<uint32_t>(),
           gst.as<uint32_t>() + (uint64_t)ngroups * PW_BINS, start, F);
}

void hash_partition_1p(pimdal_ctx* c, const uint32_t* kin, const uint32_t* vin, uint64_t n, uint32_t bits, uint2* out,
                       uint32_t* bounds, hipStream_t s) {
    PD_REQUIRE(bits >= 1 && bits <= 10, PIMDAL_E_INVALID, "one-pass partition: 1 to 10 bits");
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "one-pass partition supports up to 2^30-1 rows per call");
    const uint32_t F = 1u << bits, mask = F - 1;
    if (n == 0) {
        fill_u32(c, bounds, 0, F + 1, s);
        return;
    }
    // rows per thread (knob pw_items): 16 (16384-row tiles, one block per
    // CU) or 8 (8192-row tiles, two blocks per CU; shorter runs)
    const int64_t items = c->tune_get("pw_items", 16);
    PD_REQUIRE(items == 8 || items == 16, PIMDAL_E_INVALID, "pw_items must be 8 or 16");
    const uint32_t tile = PW_BLOCK * (uint32_t)items;
    const uint32_t ntiles = (uint32_t)ceil_div(n, tile);
    const uint32_t ngroups = (ntiles + PW_GT - 1) / PW_GT;
    const uint32_t xt = (ntiles + 7) / 8;
    // cnt[ntiles][1024] | gstatus[ngroups][1024] | ticket
    Scratch sc(c, ((uint64_t)ntiles * PW_BINS + (uint64_t)ngroups * PW_BINS + 1) * 4, s);
    uint32_t* cnt = sc.as<uint32_t>();
    uint32_t* gst = cnt + (uint64_t)ntiles * PW_BINS;
    fill_u32(c, gst, 0, (uint64_t)ngroups * PW_BINS + 1, s);
    const uint32_t cgrid = std::min<uint32_t>(ntiles, (uint32_t)c->cu_count * 4);
    if (items == 8) launch(c, "pw_count", k_pw_count<8192>, dim3(cgrid), dim3(512), 0, s, kin, n, mask, cnt);
    else launch(c, "pw_count", k_pw_count<16384>, dim3(cgrid), dim3(512), 0, s, kin, n, mask, cnt);
    launch(c, "pw_tscan", k_pw_tscan, dim3(ngroups), dim3(PW_BINS), 0, s, cnt, ntiles, gst,
           gst + (uint64_t)ngroups * PW_BINS, bounds, F);
    if (items == 8)
        launch(c, "pw_scatter", k_pw_scatter<8>, dim3(8 * xt), dim3(PW_BLOCK), 0, s, kin, vin, out, n, mask,
               (const uint32_t*)cnt, (const uint32_t*)bounds, xt);
    else
        launch(c, "pw_scatter", k_pw_scatter<16>, dim3(8 * xt), dim3(PW_BLOCK), 0, s, kin, vin, out, n, mask,
               (const uint32_t*)cnt, (const uint32_t*)bounds, xt);
}

}  // namespace pd
