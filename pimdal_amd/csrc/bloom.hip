// bloom.hip — the Bloom-filter semi-join of the PIM library
// (pimdal/join/hash_join.c:117-152 filter_kernel, 237-341 match_kernel;
// pimdal/hash/hash_steps.c:143-183 insert_bloom_filter / probe_bloom_filter).
//
// Build keys fall into 2^part_bits partitions by bits
// [shift - part_bits, shift) of hash0(key) (the MSB-first radix partition of
// hash_join.c:154-235 ends there); each partition has a 16384-bit filter in
// which hash1..hash4(key) % 16384 are set.  A probe row is kept when all four
// of its bits are set in its partition's filter.  The filter words are
// little-endian uint32, so bit b sits in byte b / 8 at bit b % 8, the byte
// layout of the reference's filters.
//
//   k_bloom_build   one pass over the build keys; with at most 32 filters
//                   (64 KB) each block sets bits in an LDS copy and ORs its
//                   non-zero words into HBM at the end (a few filters would
//                   otherwise serialize every atomic on a few hundred words);
//                   above that, device atomics straight into the filters
//   k_bloom_flags   one pass over the probe keys: a keep flag per row
//                   (filters staged in LDS when they fit, else read through
//                   L2/MALL); the kept rows' payloads and row indices are
//                   then compacted in row order by the selection pass
//   k_bloom_group   64 to 1024 filters: groups of 64 filters held in LDS,
//                   each group's blocks streaming every probe key, kept rows
//                   marked in one bitmap per group
//   k_bloom_or      keep flags from the groups' bitmaps, for the selection
#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

// Compiler-only barrier: LDS accesses of one wave execute in issue order, so
// the queue's writes and reads need ordering against code motion only.
#define PD_BG_BARRIER() __asm__ volatile("" ::: "memory")

constexpr uint32_t BLOOM_BITS = 16384;             // hash_steps.c:15 FILTER_SIZE
constexpr uint32_t BLOOM_WORDS = BLOOM_BITS / 32;  // 512 uint32 words per filter
constexpr uint32_t BLOOM_LDS_MAX_BITS = 5;         // 32 filters = 64 KB

// The shift-add-xor mixers of pimdal/hash/hash_func.h:5-58.
template <uint32_t ADD, int S1, int S2, int S3, int S4>
__device__ __forceinline__ uint32_t pim_mix(uint32_t k) {
    k += ADD;
    k += ~(k << S1);
    k ^= k >> S2;
    k += k << S3;
    k ^= k >> S4;
    return k;
}
__device__ __forceinline__ uint32_t pim_hash0(uint32_t k) { return pim_mix<272333, 5, 18, 11, 17>(k); }

// The four filter bits of a key (hash_func.h: hash_func[] = hash1..hash4).
__device__ __forceinline__ void bloom_bits(uint32_t k, uint32_t (&b)[4]) {
    b[0] = pim_mix<593627, 22, 26, 10, 28>(k) % BLOOM_BITS;
    b[1] = pim_mix<274277, 13, 20, 3, 9>(k) % BLOOM_BITS;
    b[2] = pim_mix<855467, 4, 21, 5, 16>(k) % BLOOM_BITS;
    b[3] = pim_mix<176521, 10, 25, 5, 24>(k) % BLOOM_BITS;
}

__device__ __forceinline__ uint32_t bloom_part(uint32_t k, uint32_t part_bits, uint32_t shift) {
    return part_bits ? (pim_hash0(k) >> (shift - part_bits)) & ((1u << part_bits) - 1) : 0u;
}

template <bool LDS>
__global__ __launch_bounds__(256) void k_bloom_build(const uint32_t* __restrict__ keys, uint64_t n,
                                                     uint32_t part_bits, uint32_t shift, uint32_t* __restrict__ filt) {
    __shared__ uint32_t s_f[LDS ? (BLOOM_WORDS << BLOOM_LDS_MAX_BITS) : 1];
    const uint32_t words = BLOOM_WORDS << part_bits;
    if constexpr (LDS) {
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) s_f[i] = 0;
        __syncthreads();
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t k = keys[i];
        uint32_t b[4];
        bloom_bits(k, b);
        const uint32_t base = bloom_part(k, part_bits, shift) * BLOOM_WORDS;
#pragma unroll
        for (int h = 0; h < 4; h++) {
            if constexpr (LDS)
                atomicOr(&s_f[base + (b[h] >> 5)], 1u << (b[h] & 31));
            else
                __hip_atomic_fetch_or(&filt[base + (b[h] >> 5)], 1u << (b[h] & 31), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if constexpr (LDS) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x)
            if (s_f[i]) __hip_atomic_fetch_or(&filt[i], s_f[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// LDS: 1024-thread blocks (the 64 KB filter copy allows two per CU, 32
// waves); VEC: four keys per thread per step through 16-byte loads and
// stores, so each wave has four independent lookups in flight.
template <bool LDS, bool VEC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bloom_flags(const uint32_t* __restrict__ keys, uint64_t n,
                                                       uint32_t part_bits, uint32_t shift,
                                                       const uint32_t* __restrict__ filt, uint8_t* __restrict__ flag) {
    __shared__ uint32_t s_f[LDS ? (BLOOM_WORDS << BLOOM_LDS_MAX_BITS) : 1];
    const uint32_t* f = filt;
    if constexpr (LDS) {
        const uint32_t words = BLOOM_WORDS << part_bits;
        for (uint32_t i = threadIdx.x; i < words; i += BLOCK) s_f[i] = filt[i];
        __syncthreads();
        f = s_f;
    }
    auto keep = [&](uint32_t k) -> uint32_t {
        uint32_t b[4];
        bloom_bits(k, b);
        const uint32_t base = bloom_part(k, part_bits, shift) * BLOOM_WORDS;
        uint32_t hit = 1;
#pragma unroll
        for (int h = 0; h < 4; h++) hit &= (f[base + (b[h] >> 5)] >> (b[h] & 31)) & 1u;
        return hit;
    };
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    const uint64_t n4 = VEC ? n / 4 : 0;
    for (uint64_t v = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; v < n4; v += stride) {
        const uint4 k = reinterpret_cast<const uint4*>(keys)[v];
        reinterpret_cast<uint32_t*>(flag)[v] = keep(k.x) | (keep(k.y) << 8) | (keep(k.z) << 16) | (keep(k.w) << 24);
    }
    for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i < n; i += stride)
        flag[i] = (uint8_t)keep(keys[i]);
}

// ---- probe with the filters split into groups held in LDS ----
// 64 filters (128 KB) per group: partition p is filter p % 64 of group
// p / 64.  Block b holds group b % G in LDS (with G = 8 each XCD's blocks
// hold one group) and streams the probe keys in chunks of 16384 or 32768
// rows: chunks j, j + B, ... (j = b / G) for the B = gridDim.x / G blocks of
// its group, so every group reads every key once (G reads of the key
// column in all).  Each wave tests the group of 64 rows at a time (one
// mixer), appends the rows of its own group to a per-wave LDS queue (key,
// row in the chunk and filter), and tests the four filter bits of 64 queued
// rows at a time, so the bit tests run on full waves whatever the group's
// share of the rows.  Kept rows set their bit in an LDS bitmap of the chunk,
// written out per group: bm[g][chunk * (chunk rows / 32) + word].  With the
// filters in HBM every bit test is a random L2 read (about 275 G reads/s on
// the chip); here it is an LDS read, and what remains is the G-fold read of
// the keys and the per-key group test (512 filters x 2^27 keys: 1.29 ms
// against 1.95 ms; 128 filters: 0.54 against 1.78 ms).
constexpr uint32_t BG_FILTERS = 64;
constexpr uint32_t BG_BLOCK = 1024;
constexpr uint32_t BG_WAVES = BG_BLOCK / PD_WAVE;
constexpr uint32_t BG_Q = 128;                        // queue entries per wave

// ITEMS rows per thread per chunk (chunk = 1024 ITEMS rows); the block's
// next D chunks are loading while it tests one.
template <uint32_t ITEMS, int D>
__global__ __launch_bounds__(BG_BLOCK) void k_bloom_group(const uint32_t* __restrict__ keys, uint64_t n,
                                                          uint32_t part_bits, uint32_t shift, uint32_t G,
                                                          const uint32_t* __restrict__ filt,
                                                          uint32_t* __restrict__ bm) {
    constexpr uint32_t CHUNK = BG_BLOCK * ITEMS;
    constexpr uint32_t CW = CHUNK / 32;                       // bitmap words per chunk
    constexpr uint32_t RB = ITEMS == 32 ? 15 : 14;            // bits of a row in the chunk
    static_assert((1u << RB) == CHUNK, "chunk rows");
    __shared__ __align__(16) uint32_t s_f[BG_FILTERS * BLOOM_WORDS];
    __shared__ uint32_t s_bm[CW];
    __shared__ uint32_t s_qk[BG_WAVES][BG_Q];
    __shared__ uint32_t s_qr[BG_WAVES][BG_Q];   // row in the chunk | filter in the group << RB
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint32_t g = blockIdx.x % G, B = gridDim.x / G;
    const uint32_t* fg = filt + (uint64_t)g * BG_FILTERS * BLOOM_WORDS;
    if (((uintptr_t)fg & 15) == 0) {
        for (uint32_t i = t; i < BG_FILTERS * BLOOM_WORDS / 4; i += BG_BLOCK)
            reinterpret_cast<uint4*>(s_f)[i] = reinterpret_cast<const uint4*>(fg)[i];
    } else {
        for (uint32_t i = t; i < BG_FILTERS * BLOOM_WORDS; i += BG_BLOCK) s_f[i] = fg[i];
    }
    for (uint32_t i = t; i < CW; i += BG_BLOCK) s_bm[i] = 0;
    const uint64_t nchunks = (n + CHUNK - 1) / CHUNK;
    uint32_t* out = bm + (uint64_t)g * nchunks * CW;
    auto load = [&](uint64_t ch, uint32_t (&k)[ITEMS], uint32_t& wrows) {
        const uint64_t wave_base = ch * CHUNK + (uint64_t)w * (PD_WAVE * ITEMS);
        wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
            (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * ITEMS), n - wave_base) : 0));
        const __amdgpu_buffer_rsrc_t rk = buf_rsrc(keys + wave_base, wrows * 4u);
#pragma unroll
        for (uint32_t i = 0; i < ITEMS; i++) k[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
    };
    // the queue's first entry and length (wave-uniform)
    uint32_t qh = 0, qn = 0;
    // test `cnt` queued rows from the head (lane i takes entry head + i)
    auto drain = [&](uint32_t cnt) {
        PD_BG_BARRIER();
        if (l < cnt) {
            const uint32_t e = (qh + l) & (BG_Q - 1);
            const uint32_t k = s_qk[w][e], r = s_qr[w][e];
            uint32_t b[4];
            bloom_bits(k, b);
            const uint32_t* f = s_f + (r >> RB) * BLOOM_WORDS;
            uint32_t hit = 1;
#pragma unroll
            for (int h = 0; h < 4; h++) hit &= (f[b[h] >> 5] >> (b[h] & 31)) & 1u;
            if (hit) atomicOr(&s_bm[(r & (CHUNK - 1)) >> 5], 1u << (r & 31));
        }
        PD_BG_BARRIER();
        qh = (qh + cnt) & (BG_Q - 1);
        qn -= cnt;
    };
    uint32_t key[D + 1][ITEMS], rows[D + 1];
    // this block's chunks: j0 + i B for i < K
    const uint64_t j0 = blockIdx.x / G;
    const uint64_t K = j0 < nchunks ? (nchunks - j0 + B - 1) / B : 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        rows[d] = 0;
        if ((uint64_t)d < K) load(j0 + (uint64_t)d * B, key[d], rows[d]);
    }
    __syncthreads();
    for (uint64_t i = 0; i < K; i++) {
        const uint64_t ch = j0 + i * B;
        if (i + D < K) load(j0 + (i + D) * B, key[D], rows[D]);
#pragma unroll
        for (uint32_t i = 0; i < ITEMS; i++) {
            const uint32_t part = bloom_part(key[0][i], part_bits, shift);
            const bool in = l + i * PD_WAVE < rows[0] && (part / BG_FILTERS) == g;
            const uint64_t m = ballot(in);
            if (in) {
                const uint32_t e = (qh + qn + popc64(m & lanemask_lt())) & (BG_Q - 1);
                s_qk[w][e] = key[0][i];
                s_qr[w][e] = (w * (PD_WAVE * ITEMS) + i * PD_WAVE + l) | ((part % BG_FILTERS) << RB);
            }
            qn += popc64(m);
            if (qn >= PD_WAVE) drain(PD_WAVE);
        }
        if (qn) drain(qn);
        __syncthreads();
        for (uint32_t i = t; i < CW; i += BG_BLOCK) {
            out[ch * CW + i] = s_bm[i];
            s_bm[i] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int d = 0; d < D; d++) {
#pragma unroll
            for (uint32_t i = 0; i < ITEMS; i++) key[d][i] = key[d + 1][i];
            rows[d] = rows[d + 1];
        }
    }
}

// Keep flags from the groups' bitmaps: flag[row] = 1 when row's bit is set
// in any group's bitmap (each row's bit is set in its own group's only).
// A wave takes 64 bitmap words (2048 rows): lane l ORs word l over the
// groups, then the wave writes the 2048 one-byte flags as 2 rounds of 64
// contiguous 16-byte stores, each lane taking its 16 rows' bits from the
// lane that holds their word.
__device__ __forceinline__ uint32_t bloom_spread4(uint32_t nib) {   // 4 bits -> 4 bytes of 0 / 1
    return (nib * 0x204081u) & 0x01010101u;
}

__global__ __launch_bounds__(256) void k_bloom_or(const uint32_t* __restrict__ bm, uint32_t G, uint64_t gstride,
                                                  uint64_t n, uint8_t* __restrict__ flag) {
    const uint32_t l = lane();
    const uint64_t nw = (n + 31) / 32;
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / PD_WAVE;
    const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / PD_WAVE;
    const bool vec = ((uintptr_t)flag % 16) == 0;
    for (uint64_t w0 = wave * PD_WAVE; w0 < nw; w0 += nwaves * PD_WAVE) {
        uint32_t x = 0;
        if (w0 + l < nw)
            for (uint32_t gg = 0; gg < G; gg++) x |= bm[gg * gstride + w0 + l];
#pragma unroll
        for (uint32_t q = 0; q < 2; q++) {
            const uint32_t j = q * PD_WAVE + l;               // 16-row group of this wave
            const uint32_t xw = (uint32_t)__shfl((int)x, (int)(j / 2), PD_WAVE);
            const uint32_t h = (xw >> ((j % 2) * 16)) & 0xFFFFu;
            const uint64_t r0 = w0 * 32 + (uint64_t)j * 16;
            if (vec && r0 + 16 <= n) {
                reinterpret_cast<uint4*>(flag + r0)[0] = make_uint4(bloom_spread4(h & 15u), bloom_spread4((h >> 4) & 15u),
                                                                    bloom_spread4((h >> 8) & 15u), bloom_spread4(h >> 12));
            } else {
                for (uint32_t i = 0; i < 16 && r0 + i < n; i++) flag[r0 + i] = (uint8_t)((h >> i) & 1u);
            }
        }
    }
}

static void check_bloom_args(uint32_t part_bits, uint32_t shift) {
    PD_REQUIRE(part_bits <= 16, PIMDAL_E_INVALID, "part_bits above 16 (filters of 2^part_bits x 2 KB)");
    PD_REQUIRE(shift <= 32 && shift >= part_bits, PIMDAL_E_INVALID,
               "shift must lie in [part_bits, 32]: partitions are bits [shift - part_bits, shift) of hash0");
}

static void bloom_build(pimdal_ctx* c, const uint32_t* keys, uint64_t n, uint32_t part_bits, uint32_t shift,
                        uint32_t* filt, hipStream_t s) {
    fill_u32(c, filt, 0, (uint64_t)BLOOM_WORDS << part_bits, s);
    if (n == 0) return;
    const bool lds = part_bits <= BLOOM_LDS_MAX_BITS;
    // LDS copies: one block per CU (each block ORs its whole table out once)
    const uint64_t cap = lds ? (uint64_t)c->cu_count : (uint64_t)c->cu_count * 8;
    const dim3 grid((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), cap)));
    if (lds)
        launch(c, "bloom_build", k_bloom_build<true>, grid, dim3(256), 0, s, keys, n, part_bits, shift, filt);
    else
        launch(c, "bloom_build", k_bloom_build<false>, grid, dim3(256), 0, s, keys, n, part_bits, shift, filt);
}

static void bloom_probe(pimdal_ctx* c, const uint32_t* keys, const uint32_t* pay, uint64_t n, uint32_t part_bits,
                        uint32_t shift, const uint32_t* filt, uint32_t* out_pay, uint32_t* out_row, uint64_t* d_count,
                        hipStream_t s) {
    if (n == 0) {
        fill_u64(c, d_count, 0, 1, s);
        return;
    }
    Scratch flags(c, n, s);   // one byte per row
    const bool lds = part_bits <= BLOOM_LDS_MAX_BITS;
    // filter groups in LDS (knob bloom_groups, default on): 64 to 1024
    // filters, probe sides of at least bloom_group_min_rows rows
    if (!lds && part_bits <= 10 && c->tune_get("bloom_groups", 1) != 0 &&
        n >= (uint64_t)c->tune_get("bloom_group_min_rows", 1 << 16)) {
        const uint32_t G = 1u << (part_bits - 6);
        // chunk shape (knob bloom_group_shape): 1 = 32768-row chunks, the
        // next one loading (default; 1.29 ms for 512 filters x 2^27 keys);
        // 0 = 16384-row chunks, next loading (1.39 ms); 2 = 16384-row
        // chunks, the next two loading (1.31 ms)
        const int64_t shape = c->tune_get("bloom_group_shape", 1);
        PD_REQUIRE(shape >= 0 && shape <= 2, PIMDAL_E_INVALID, "bloom_group_shape must be 0, 1 or 2");
        const uint64_t chunk = shape == 1 ? 32768 : 16384;
        const uint64_t nchunks = ceil_div(n, chunk);
        const uint64_t gstride = nchunks * (chunk / 32);
        Scratch bm(c, (uint64_t)G * gstride * 4, s);
        const uint32_t grid = std::max<uint32_t>(G, ((uint32_t)c->cu_count / G) * G);
#define PD_BG(IT, D)                                                                                             \
    launch(c, "bloom_group", k_bloom_group<IT, D>, dim3(grid), dim3(BG_BLOCK), 0, s, keys, n, part_bits, shift, G, \
           filt, bm.as<uint32_t>())
        if (shape == 1) PD_BG(32, 1);
        else if (shape == 2) PD_BG(16, 2);
        else PD_BG(16, 1);
#undef PD_BG
        const uint64_t nw = ceil_div(n, 32);   // a wave per 64 bitmap words, four waves per block
        launch(c, "bloom_or", k_bloom_or, dim3((uint32_t)std::min<uint64_t>(ceil_div(nw, 256), 16384)), dim3(256), 0, s,
               (const uint32_t*)bm.as<uint32_t>(), G, gstride, n, flags.as<uint8_t>());
        select_range(c, PIMDAL_U8, flags.get(), pay ? pay : keys, n, 1, 2, out_pay, out_row, d_count, s);
        return;
    }
    const bool vec = (uintptr_t)keys % 16 == 0;   // the flags are pool memory, 256-byte aligned
    const uint64_t work = vec ? ceil_div(n, 4) : n;
    // LDS: two 1024-thread blocks per CU, each copying the filters once;
    // HBM filters: 256-thread blocks, eight per CU
    const int block = lds ? 1024 : 256;
    const uint64_t cap = (uint64_t)c->cu_count * (lds ? 2 : 8);
    const dim3 grid((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(work, block), cap)));
#define PD_BLOOM_FLAGS(L, V, B)                                                                                \
    launch(c, "bloom_flags", k_bloom_flags<L, V, B>, grid, dim3(B), 0, s, keys, n, part_bits, shift, filt,    \
           flags.as<uint8_t>())
    if (lds && vec) PD_BLOOM_FLAGS(true, true, 1024);
    else if (lds) PD_BLOOM_FLAGS(true, false, 1024);
    else if (vec) PD_BLOOM_FLAGS(false, true, 256);
    else PD_BLOOM_FLAGS(false, false, 256);
#undef PD_BLOOM_FLAGS
    // kept rows (flag in [1, 2)) in row order: payload (the key when there is none) and row index
    select_range(c, PIMDAL_U8, flags.get(), pay ? pay : keys, n, 1, 2, out_pay, out_row, d_count, s);
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_bloom_build_dev(pimdal_ctx* c, const uint32_t* keys, uint64_t n, uint32_t part_bits,
                                      uint32_t shift, uint32_t* filters) {
    PD_API_BEGIN
    PD_REQUIRE(c && filters && (keys || n == 0), PIMDAL_E_INVALID, "NULL argument");
    check_bloom_args(part_bits, shift);
    bloom_build(c, keys, n, part_bits, shift, filters, c->stream);
    PD_API_END
}

extern "C" int pimdal_bloom_probe_dev(pimdal_ctx* c, const uint32_t* keys, const uint32_t* payload, uint64_t n,
                                      uint32_t part_bits, uint32_t shift, const uint32_t* filters,
                                      uint32_t* out_payload, uint32_t* out_rowid, uint64_t* d_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && filters && d_count && (keys || n == 0), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(n < (1ull << 32), PIMDAL_E_INVALID, "probe side above 2^32-1 rows (uint32 row ids)");
    check_bloom_args(part_bits, shift);
    bloom_probe(c, keys, payload, n, part_bits, shift, filters, out_payload, out_rowid, d_count, c->stream);
    PD_API_END
}

extern "C" int pimdal_bloom_semi_join_dev(pimdal_ctx* c, const uint32_t* build_keys, uint64_t nb,
                                          const uint32_t* probe_keys, const uint32_t* probe_payload, uint64_t np,
                                          uint32_t part_bits, uint32_t shift, uint32_t* out_payload,
                                          uint32_t* out_rowid, uint64_t* d_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && d_count && (build_keys || nb == 0) && (probe_keys || np == 0), PIMDAL_E_INVALID,
               "NULL argument");
    PD_REQUIRE(np < (1ull << 32), PIMDAL_E_INVALID, "probe side above 2^32-1 rows (uint32 row ids)");
    check_bloom_args(part_bits, shift);
    Scratch filt(c, ((uint64_t)BLOOM_WORDS << part_bits) * 4, c->stream);
    bloom_build(c, build_keys, nb, part_bits, shift, filt.as<uint32_t>(), c->stream);
    bloom_probe(c, probe_keys, probe_payload, np, part_bits, shift, filt.as<uint32_t>(), out_payload, out_rowid,
                d_count, c->stream);
    PD_API_END
}
