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
#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

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
                                                       const uint32_t* __restrict__ filt, uint32_t* __restrict__ flag) {
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
        uint4 o;
        o.x = keep(k.x);
        o.y = keep(k.y);
        o.z = keep(k.z);
        o.w = keep(k.w);
        reinterpret_cast<uint4*>(flag)[v] = o;
    }
    for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i < n; i += stride) flag[i] = keep(keys[i]);
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
    Scratch flags(c, n * 4, s);
    const bool lds = part_bits <= BLOOM_LDS_MAX_BITS;
    const bool vec = (uintptr_t)keys % 16 == 0;   // the flags are pool memory, 256-byte aligned
    const uint64_t work = vec ? ceil_div(n, 4) : n;
    // LDS: two 1024-thread blocks per CU, each copying the filters once;
    // HBM filters: 256-thread blocks, eight per CU
    const int block = lds ? 1024 : 256;
    const uint64_t cap = (uint64_t)c->cu_count * (lds ? 2 : 8);
    const dim3 grid((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(work, block), cap)));
#define PD_BLOOM_FLAGS(L, V, B)                                                                                \
    launch(c, "bloom_flags", k_bloom_flags<L, V, B>, grid, dim3(B), 0, s, keys, n, part_bits, shift, filt,    \
           flags.as<uint32_t>())
    if (lds && vec) PD_BLOOM_FLAGS(true, true, 1024);
    else if (lds) PD_BLOOM_FLAGS(true, false, 1024);
    else if (vec) PD_BLOOM_FLAGS(false, true, 256);
    else PD_BLOOM_FLAGS(false, false, 256);
#undef PD_BLOOM_FLAGS
    // kept rows (flag in [1, 2)) in row order: payload (the key when there is none) and row index
    select_range(c, PIMDAL_U32, flags.get(), pay ? pay : keys, n, 1, 2, out_pay, out_row, d_count, s);
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
