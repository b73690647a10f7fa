// hash_join.hip — radix-partitioned hash join with LDS-resident tables.
//
// Reference semantics: reference/db_operators/hash_join.cpp:57-219
// (two-pass partition by hash % part_n, then per bucket a hash-set build
// and probe) and the PIM path pimdal/join/hash_join.c:40-235, 343-430
// (multi-level radix partition, 256-slot linear-probing tables per
// partition, probe emitting (outer ptr, inner ptr)).  Output here is one
// (probe payload, build payload) row per matching pair, in no particular
// order (compare after a canonical sort, as the reference's own order
// depends on its thread schedule).
//
// Plan on gfx950:
//  1. radix-partition both relations by the reference hash
//     (hash_join.cpp:20-25) into F partitions: per-block LDS histograms
//     (k_part_hist), one device scan over [partition][block], then a
//     scatter that re-reads the rows and claims slots with LDS counters
//     (k_part_scatter), writing (key, payload) pairs;
//  2. split every probe partition into chunks of at most CH rows; each
//     chunk is one work item, so a hot partition under Zipf skew becomes
//     many blocks instead of one straggler;
//  3. per work item (k_hj_join): build the partition's table in LDS
//     (linear probing, occupancy bits kept apart from the keys so every
//     32-bit key value is legal), then stream the probe chunk through it.
//     The first launch assumes unique build keys and emits at most one
//     match per probe row; it also counts every match, so duplicate build
//     keys are detected and the join is re-run counting then writing all
//     pairs.
#include <algorithm>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int PT_BLOCK = 512;
constexpr int HJ_BLOCK = 512;
constexpr int HJ_TS = 8192;           // LDS table slots (8 B each)
constexpr int HJ_ROWS = 4;            // probe rows per thread per round

// ---------------- radix partition of (key, payload) rows ----------------

__global__ __launch_bounds__(PT_BLOCK) void k_part_hist(const uint32_t* __restrict__ keys, uint64_t n, uint64_t chunk,
                                                        uint32_t F, uint32_t* __restrict__ H) {
    extern __shared__ uint32_t h[];
    const uint32_t t = threadIdx.x, b = blockIdx.x, G = gridDim.x;
    for (uint32_t f = t; f < F; f += PT_BLOCK) h[f] = 0;
    __syncthreads();
    const uint64_t start = (uint64_t)b * chunk, end = min(n, start + chunk);
    const uint32_t fm = F - 1;
    for (uint64_t i = start + t; i < end; i += PT_BLOCK) atomicAdd(&h[pd::ref_hash(keys[i]) & fm], 1u);
    __syncthreads();
    for (uint32_t f = t; f < F; f += PT_BLOCK) H[(uint64_t)f * G + b] = h[f];
}

__global__ __launch_bounds__(PT_BLOCK) void k_part_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ pay, uint64_t n,
                                                           uint64_t chunk, uint32_t F,
                                                           const uint32_t* __restrict__ Hs, uint2* __restrict__ out) {
    extern __shared__ uint32_t ctr[];
    const uint32_t t = threadIdx.x, b = blockIdx.x, G = gridDim.x;
    for (uint32_t f = t; f < F; f += PT_BLOCK) ctr[f] = Hs[(uint64_t)f * G + b];
    __syncthreads();
    const uint64_t start = (uint64_t)b * chunk, end = min(n, start + chunk);
    const uint32_t fm = F - 1;
    for (uint64_t i = start + t; i < end; i += PT_BLOCK) {
        const uint32_t k = keys[i];
        const uint32_t p = pay ? pay[i] : k;
        const uint32_t pos = atomicAdd(&ctr[pd::ref_hash(k) & fm], 1u);
        if (pos < n) out[pos] = make_uint2(k, p);
    }
}

// Partitioned relation: rows in `rows`, partition f spans
// [H[f*G], H[(f+1)*G]) where H is the scanned [partition][block] array.
struct Parted {
    uint2* rows;
    uint32_t* H;
    uint32_t G;
};

static uint32_t part_blocks(pimdal_ctx* c, uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 16384), (uint64_t)c->cu_count * 2));
}

void radix_partition(pimdal_ctx* c, const uint32_t* keys, const uint32_t* pay, uint64_t n, uint32_t F, uint2* out,
                     uint32_t* H, uint32_t G, hipStream_t s) {
    const uint64_t chunk = ceil_div(n, G);
    const uint32_t shm = F * 4;
    launch(c, "part_hist", k_part_hist, dim3(G), dim3(PT_BLOCK), shm, s, keys, n, chunk, F, H);
    exclusive_scan_u32(c, H, H, (uint64_t)F * G, nullptr, s);
    launch(c, "part_scatter", k_part_scatter, dim3(G), dim3(PT_BLOCK), shm, s, keys, pay, n, chunk, F,
           (const uint32_t*)H, out);
}

// ---------------- work items over probe partitions ----------------

__global__ void k_hj_chunks(const uint32_t* __restrict__ HS, uint32_t GS, uint32_t F, uint32_t CH,
                            uint32_t* __restrict__ nch) {
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        const uint32_t len = HS[(uint64_t)(f + 1) * GS] - HS[(uint64_t)f * GS];
        nch[f] = (len + CH - 1) / CH;
    }
}

// flags[0]: a probe row matched more than one build row
// flags[1]: a build partition exceeded the LDS table
// flags[2..3]: total emitted pairs (uint64)
enum { HJ_UNIQUE = 0, HJ_COUNT = 1, HJ_WRITE = 2 };

template <int PHASE>
__global__ __launch_bounds__(HJ_BLOCK, 2) void k_hj_join(const uint2* __restrict__ R, const uint32_t* __restrict__ HR,
                                                         uint32_t GR, const uint2* __restrict__ S,
                                                         const uint32_t* __restrict__ HS, uint32_t GS, uint32_t F,
                                                         uint32_t CH, const uint32_t* __restrict__ wstart,
                                                         uint32_t* __restrict__ wcount,
                                                         const uint32_t* __restrict__ woff,
                                                         uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                         uint64_t out_cap, uint32_t* __restrict__ flags) {
    __shared__ uint2 tab[HJ_TS];
    __shared__ uint32_t occ[HJ_TS / 32];
    __shared__ uint32_t s_wsum[HJ_BLOCK / PD_WAVE];
    __shared__ uint32_t s_f, s_bad;

    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (t == 0) {
        // partition of this work item: last f with wstart[f] <= wi
        uint32_t lo = 0, hi = F;  // answer in [lo, hi)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_f = lo;
        s_bad = 0;
    }
    for (uint32_t i = t; i < HJ_TS / 32; i += HJ_BLOCK) occ[i] = 0;
    __syncthreads();
    const uint32_t f = s_f;
    const uint32_t ck = wi - wstart[f];
    const uint64_t r0 = HR[(uint64_t)f * GR], r1 = HR[(uint64_t)(f + 1) * GR];
    const uint64_t sp0 = HS[(uint64_t)f * GS], sp1 = HS[(uint64_t)(f + 1) * GS];
    const uint64_t s0 = sp0 + (uint64_t)ck * CH, s1 = min(sp1, s0 + CH);
    if (r1 - r0 > (uint64_t)HJ_TS * 7 / 8) {
        if (t == 0) {
            atomicOr(&flags[1], 1u);
            if (PHASE != HJ_WRITE) wcount[wi] = 0;
        }
        return;
    }

    // ---- build ----
    for (uint64_t i = r0 + t; i < r1; i += HJ_BLOCK) {
        const uint2 kv = R[i];
        uint32_t h = pd::slot_hash(kv.x) & (HJ_TS - 1);
        while (true) {
            const uint32_t bit = 1u << (h & 31);
            const uint32_t prev = atomicOr(&occ[h >> 5], bit);
            if (!(prev & bit)) {
                tab[h] = kv;
                break;
            }
            h = (h + 1) & (HJ_TS - 1);
        }
    }
    __syncthreads();

    // ---- probe in rounds of HJ_BLOCK * HJ_ROWS rows ----
    uint64_t written = 0;        // block-uniform running output count
    const uint64_t wbase = (PHASE == HJ_WRITE) ? woff[wi] : s0;
    bool dup = false;
    for (uint64_t rb = s0; rb < s1; rb += (uint64_t)HJ_BLOCK * HJ_ROWS) {
        uint32_t key[HJ_ROWS], pay[HJ_ROWS], first[HJ_ROWS], m[HJ_ROWS];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < HJ_ROWS; k++) {
            const uint64_t i = rb + (uint64_t)k * HJ_BLOCK + t;
            m[k] = 0;
            first[k] = 0;
            if (i < s1) {
                const uint2 kv = S[i];
                key[k] = kv.x;
                pay[k] = kv.y;
                uint32_t h = pd::slot_hash(kv.x) & (HJ_TS - 1);
                while (occ[h >> 5] & (1u << (h & 31))) {
                    const uint2 e = tab[h];
                    if (e.x == kv.x) {
                        if (m[k] == 0) first[k] = e.y;
                        m[k]++;
                    }
                    h = (h + 1) & (HJ_TS - 1);
                }
            }
            if (PHASE == HJ_UNIQUE) {
                dup |= m[k] > 1;
                cnt += m[k] ? 1 : 0;
            } else {
                cnt += m[k];
            }
        }
        // block exclusive scan of cnt
        const uint32_t inc = pd::wave_incl_scan(cnt);
        if (l == PD_WAVE - 1) s_wsum[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < HJ_BLOCK / PD_WAVE; i++) {
            const uint32_t v = s_wsum[i];
            pre += (i < (int)w) ? v : 0;
            tot += v;
        }
        __syncthreads();
        if (PHASE != HJ_COUNT) {
            uint64_t o = wbase + written + pre + inc - cnt;
#pragma unroll
            for (int k = 0; k < HJ_ROWS; k++) {
                if (!m[k]) continue;
                if (PHASE == HJ_UNIQUE) {
                    if (o < out_cap) { out_p[o] = pay[k]; out_b[o] = first[k]; }
                    o++;
                } else {
                    uint32_t h = pd::slot_hash(key[k]) & (HJ_TS - 1);
                    while (occ[h >> 5] & (1u << (h & 31))) {
                        const uint2 e = tab[h];
                        if (e.x == key[k]) {
                            if (o < out_cap) { out_p[o] = pay[k]; out_b[o] = e.y; }
                            o++;
                        }
                        h = (h + 1) & (HJ_TS - 1);
                    }
                }
            }
        }
        written += tot;
    }
    if (PHASE == HJ_UNIQUE && dup) s_bad = 1;
    __syncthreads();
    if (t == 0) {
        if (PHASE != HJ_WRITE) wcount[wi] = (uint32_t)written;
        if (PHASE == HJ_UNIQUE && s_bad) atomicOr(&flags[0], 1u);
        if (written) atomicAdd(reinterpret_cast<unsigned long long*>(flags + 2), (unsigned long long)written);
    }
}

// Moves each work item's compacted output [wbase(wi), +wcount) to woff[wi].
__global__ __launch_bounds__(256) void k_hj_compact(const uint32_t* __restrict__ HS, uint32_t GS, uint32_t F,
                                                    uint32_t CH, const uint32_t* __restrict__ wstart,
                                                    const uint32_t* __restrict__ wcount,
                                                    const uint32_t* __restrict__ woff,
                                                    const uint32_t* __restrict__ in_p, const uint32_t* __restrict__ in_b,
                                                    uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                    uint64_t out_cap) {
    __shared__ uint32_t s_f;
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (threadIdx.x == 0) {
        uint32_t lo = 0, hi = F;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_f = lo;
    }
    __syncthreads();
    const uint32_t f = s_f;
    const uint64_t src = HS[(uint64_t)f * GS] + (uint64_t)(wi - wstart[f]) * CH;
    const uint64_t dst = woff[wi];
    const uint32_t cnt = wcount[wi];
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        if (dst + j < out_cap) {
            out_p[dst + j] = in_p[src + j];
            out_b[dst + j] = in_b[src + j];
        }
    }
}

static uint32_t next_pow2(uint64_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Returns the number of matching pairs; the first min(count, cap) are in
// out_p / out_b.
uint64_t hash_join(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                   const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap, hipStream_t s) {
    PD_REQUIRE(nb < (1ull << 32) && np < (1ull << 32), PIMDAL_E_INVALID, "join inputs must have < 2^32 rows");
    PD_REQUIRE((bk || !nb) && (pk || !np), PIMDAL_E_INVALID, "key pointer is NULL");
    if (nb == 0 || np == 0) return 0;
    const int64_t tuned = c->tune_get("hj_fanout", 0);
    uint32_t F = tuned > 0 ? next_pow2((uint64_t)tuned)
                           : next_pow2(std::max<uint64_t>(1, ceil_div(nb, (uint64_t)c->tune_get("hj_rows_per_part", 4096))));
    const uint32_t CH = (uint32_t)c->tune_get("hj_chunk", 32768);
    const uint32_t GR = part_blocks(c, nb), GS = part_blocks(c, np);
    Scratch rrows(c, nb * 8, s), srows(c, np * 8, s);
    Scratch flags(c, 16, s);
    const bool direct = cap >= np && out_p && out_b;
    Scratch tmp_p(c, direct ? 0 : np * 4, s), tmp_b(c, direct ? 0 : np * 4, s);
    uint32_t* p0 = direct ? out_p : tmp_p.as<uint32_t>();
    uint32_t* b0 = direct ? out_b : tmp_b.as<uint32_t>();
    for (int attempt = 0;; attempt++) {
        PD_REQUIRE(F <= 16384, PIMDAL_E_UNSUPPORTED,
                   "hash join: a build partition does not fit the LDS table (heavy duplicate build keys); "
                   "use the sort-merge join");
        Scratch HR(c, ((uint64_t)F * GR + 1) * 4, s), HS(c, ((uint64_t)F * GS + 1) * 4, s);
        radix_partition(c, bk, bp, nb, F, rrows.as<uint2>(), HR.as<uint32_t>(), GR, s);
        radix_partition(c, pk, pp, np, F, srows.as<uint2>(), HS.as<uint32_t>(), GS, s);
        Scratch wst(c, ((uint64_t)F + 1) * 4, s);
        launch(c, "hj_chunks", k_hj_chunks, dim3(std::min<uint32_t>(ceil_div(F, 256), 1024)), dim3(256), 0, s,
               (const uint32_t*)HS.as<uint32_t>(), GS, F, CH, wst.as<uint32_t>());
        exclusive_scan_u32(c, wst.as<uint32_t>(), wst.as<uint32_t>(), F, nullptr, s);
        const uint64_t max_wi = F + ceil_div(np, CH);
        Scratch wcnt(c, (max_wi + 1) * 4, s), woff(c, (max_wi + 1) * 4, s);
        fill_u32(c, flags.get(), 0, 4, s);
        fill_u32(c, wcnt.get(), 0, max_wi + 1, s);
        launch(c, "hj_join", k_hj_join<HJ_UNIQUE>, dim3((uint32_t)max_wi), dim3(HJ_BLOCK), 0, s,
               (const uint2*)rrows.as<uint2>(), (const uint32_t*)HR.as<uint32_t>(), GR,
               (const uint2*)srows.as<uint2>(), (const uint32_t*)HS.as<uint32_t>(), GS, F, CH,
               (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), (const uint32_t*)nullptr, p0, b0,
               (uint64_t)np, flags.as<uint32_t>());
        uint32_t fl[4];
        PD_HIP(hipMemcpyAsync(fl, flags.get(), 16, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        if (fl[1]) {  // a build partition overflowed the table: more partitions
            F *= 2;
            continue;
        }
        uint64_t total = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32);
        if (fl[0]) {
            // duplicate build keys: count, scan, write every pair
            fill_u32(c, wcnt.get(), 0, max_wi + 1, s);
            fill_u32(c, flags.get(), 0, 4, s);
            launch(c, "hj_join_count", k_hj_join<HJ_COUNT>, dim3((uint32_t)max_wi), dim3(HJ_BLOCK), 0, s,
                   (const uint2*)rrows.as<uint2>(), (const uint32_t*)HR.as<uint32_t>(), GR,
                   (const uint2*)srows.as<uint2>(), (const uint32_t*)HS.as<uint32_t>(), GS, F, CH,
                   (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), (const uint32_t*)nullptr,
                   (uint32_t*)nullptr, (uint32_t*)nullptr, (uint64_t)0, flags.as<uint32_t>());
            exclusive_scan_u32(c, wcnt.as<uint32_t>(), woff.as<uint32_t>(), max_wi, nullptr, s);
            PD_HIP(hipMemcpyAsync(fl, flags.get(), 16, hipMemcpyDeviceToHost, s));
            PD_HIP(hipStreamSynchronize(s));
            total = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32);
            PD_REQUIRE(total < (1ull << 32), PIMDAL_E_UNSUPPORTED, "join output exceeds 2^32-1 pairs");
            if (out_p && out_b && cap)
                launch(c, "hj_join_write", k_hj_join<HJ_WRITE>, dim3((uint32_t)max_wi), dim3(HJ_BLOCK), 0, s,
                       (const uint2*)rrows.as<uint2>(), (const uint32_t*)HR.as<uint32_t>(), GR,
                       (const uint2*)srows.as<uint2>(), (const uint32_t*)HS.as<uint32_t>(), GS, F, CH,
                       (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(),
                       (const uint32_t*)woff.as<uint32_t>(), out_p, out_b, cap, flags.as<uint32_t>());
            return total;
        }
        if (direct && total == np) return total;  // every probe row matched once: already dense
        // compact the per-work-item runs
        exclusive_scan_u32(c, wcnt.as<uint32_t>(), woff.as<uint32_t>(), max_wi, nullptr, s);
        if (out_p && out_b && cap) {
            if (direct) {
                Scratch cp(c, total * 4 + 4, s), cb(c, total * 4 + 4, s);
                launch(c, "hj_compact", k_hj_compact, dim3((uint32_t)max_wi), dim3(256), 0, s,
                       (const uint32_t*)HS.as<uint32_t>(), GS, F, CH, (const uint32_t*)wst.as<uint32_t>(),
                       (const uint32_t*)wcnt.as<uint32_t>(), (const uint32_t*)woff.as<uint32_t>(),
                       (const uint32_t*)p0, (const uint32_t*)b0, cp.as<uint32_t>(), cb.as<uint32_t>(), total);
                copy_d2d(c, out_p, cp.get(), std::min(total, cap) * 4, s);
                copy_d2d(c, out_b, cb.get(), std::min(total, cap) * 4, s);
                PD_HIP(hipStreamSynchronize(s));
            } else {
                launch(c, "hj_compact", k_hj_compact, dim3((uint32_t)max_wi), dim3(256), 0, s,
                       (const uint32_t*)HS.as<uint32_t>(), GS, F, CH, (const uint32_t*)wst.as<uint32_t>(),
                       (const uint32_t*)wcnt.as<uint32_t>(), (const uint32_t*)woff.as<uint32_t>(),
                       (const uint32_t*)p0, (const uint32_t*)b0, out_p, out_b, cap);
            }
        }
        return total;
    }
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_hash_join_dev(pimdal_ctx* c, const uint32_t* build_key, const uint32_t* build_payload,
                                    uint64_t n_build, const uint32_t* probe_key, const uint32_t* probe_payload,
                                    uint64_t n_probe, uint32_t* out_probe_payload, uint32_t* out_build_payload,
                                    uint64_t out_capacity, uint64_t* out_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && out_count, PIMDAL_E_INVALID, "NULL argument");
    const uint64_t k = hash_join(c, build_key, build_payload, n_build, probe_key, probe_payload, n_probe,
                                 out_probe_payload, out_build_payload, out_capacity, c->stream);
    PD_HIP(hipStreamSynchronize(c->stream));
    *out_count = k;
    PD_REQUIRE(k <= out_capacity, PIMDAL_E_CAPACITY, "join output larger than out_capacity (count returned)");
    PD_API_END
}
