// merge_join.hip — sort-merge join.
//
// Reference semantics: reference/db_operators/sort_merge_join.cpp:48-103
// (sort inner, sort outer, then merge with a monotone inner cursor per
// 8000-row block) and pimdal/join/sort_merge.c:54-152 (per-tasklet merge of
// sorted blocks, matches compacted in order).  The reference merge shares
// one cursor across threads and reads the inner size from the outer table;
// this version is the corrected join: every matching pair, in
// (key, probe row, build row) order.
//
// Plan on gfx950: both sides go through the onesweep radix sort (bits that
// never vary are skipped, so 24-bit keys take 3 passes).  If the sorted
// build keys are unique, one kernel merges: each 4096-row tile of the
// sorted probe side finds the build range it can match with two binary
// searches, caches that range's keys in LDS, searches each probe key there,
// and writes the matches at offsets from a decoupled look-back, so the
// output is in probe order.  With repeated build keys: per probe row the match range
// [lower, upper), a device scan of the range lengths, then a write pass.
#include <algorithm>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int SM_BLOCK = 256;
constexpr int SM_ITEMS = 16;
constexpr int SM_TILE = SM_BLOCK * SM_ITEMS;  // 4096
constexpr int SM_WAVES = SM_BLOCK / PD_WAVE;
constexpr uint32_t SM_RCAP = 2048;            // build keys cached per tile

__device__ __forceinline__ uint64_t lower_bound_u32(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t x) {
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t upper_bound_u32(const uint32_t* a, uint64_t lo, uint64_t hi, uint32_t x) {
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void k_adjacent_dup(const uint32_t* __restrict__ k, uint64_t n, uint32_t* __restrict__ flag) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x + 1; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (k[i] == k[i - 1]) { *flag = 1; return; }
}

// Unique build keys: one 4096-row tile of the sorted probe side per block.
// Probe keys and payloads are loaded while the tile's build-key range is
// cached in LDS; each key is searched there (binary search in HBM when the
// range exceeds the cache); build payloads are gathered before the
// look-back; matches are written straight from registers in
// (item, wave, lane) = probe-row order, so every store instruction writes
// consecutive output slots.
__global__ __launch_bounds__(SM_BLOCK) void k_smj_unique(const uint32_t* __restrict__ rk, const uint32_t* __restrict__ rp,
                                                         uint64_t nr, const uint32_t* __restrict__ sk,
                                                         const uint32_t* __restrict__ sp, uint64_t ns,
                                                         uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                         uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
                                                         uint64_t ntiles, uint64_t* __restrict__ d_count,
                                                         const uint64_t* __restrict__ ranges) {
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_rk[SM_RCAP];
    __shared__ uint32_t s_cnt[SM_ITEMS][SM_WAVES];

    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t base = tile * SM_TILE;
    const uint64_t end = min(ns, base + SM_TILE);
    // probe rows base + i*BLOCK + t: row order is (i, wave, lane)
    uint32_t key[SM_ITEMS], pay[SM_ITEMS];
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const uint64_t j = base + (uint64_t)i * SM_BLOCK + t;
        key[i] = j < end ? sk[j] : 0u;
        pay[i] = j < end ? sp[j] : 0u;
    }
    const uint64_t lo = ranges[2 * tile], hi = ranges[2 * tile + 1];
    const bool cached = (hi - lo) <= SM_RCAP;
    if (cached)
        for (uint64_t i = lo + t; i < hi; i += SM_BLOCK) s_rk[i - lo] = rk[i];
    __syncthreads();

    uint32_t bval[SM_ITEMS];
    uint32_t hit = 0;
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const uint64_t j = base + (uint64_t)i * SM_BLOCK + t;
        bool m = false;
        bval[i] = 0;
        if (j < end) {
            uint64_t p;
            if (cached) {
                uint32_t a = 0, b = (uint32_t)(hi - lo);
                while (a < b) {
                    const uint32_t mid = (a + b) >> 1;
                    if (s_rk[mid] < key[i]) a = mid + 1; else b = mid;
                }
                m = a < (uint32_t)(hi - lo) && s_rk[a] == key[i];
                p = lo + a;
            } else {
                p = lower_bound_u32(rk, lo, hi, key[i]);
                m = p < hi && rk[p] == key[i];
            }
            if (m) bval[i] = rp[p];
        }
        const uint64_t bal = pd::ballot(m);
        hit |= (uint32_t)m << i;
        if (l == 0) s_cnt[i][w] = pd::popc64(bal);
    }
    __syncthreads();
    // exclusive scan over (item, wave) counts: 64 entries, one wave
    if (w == 0) {
        const uint32_t v = (&s_cnt[0][0])[l];
        const uint32_t inc = pd::wave_incl_scan(v);
        (&s_cnt[0][0])[l] = inc - v;
        const uint32_t total = __shfl(inc, PD_WAVE - 1, PD_WAVE);
        const uint64_t e = pd::lookback_single<4>(status, tile, total);
        if (l == 0) {
            s_excl = e;
            if (tile == ntiles - 1) *d_count = e + total;
        }
    }
    __syncthreads();
    const uint64_t excl = s_excl;
    const uint64_t lt = pd::lanemask_lt();
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const bool m = (hit >> i) & 1;
        const uint64_t bal = pd::ballot(m);
        if (m) {
            const uint64_t o = excl + s_cnt[i][w] + pd::popc64(bal & lt);
            if (o < ns) {
                out_p[o] = pay[i];
                out_b[o] = bval[i];
            }
        }
    }
}

// Unique build keys, direct output slots: probe row j (sorted order) writes
// its pair to slot j when it matches, and valid[j / 64] bit j % 64 records
// that it did; tcnt[tile] = the tile's matches.  With every probe row
// matched (a foreign-key join) the output is complete, in probe order, as
// written; otherwise k_smj_pack closes the holes (order kept).  No tile
// waits on another (no ticket, no look-back).  Matching is a merge: each
// thread takes 16 consecutive probe rows of the tile (staged in LDS, one pad
// word per 16 so a wave's rows fall in distinct banks), binary-searches the
// build keys for its first row and walks forward for the rest.
constexpr uint32_t SM_MISS = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t smj_pad(uint32_t r) { return r + (r >> 4); }

__global__ __launch_bounds__(SM_BLOCK) void k_smj_direct(const uint32_t* __restrict__ rk, const uint32_t* __restrict__ rp,
                                                         const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sp,
                                                         uint64_t ns, uint32_t* __restrict__ out_p,
                                                         uint32_t* __restrict__ out_b, uint64_t* __restrict__ valid,
                                                         uint32_t* __restrict__ tcnt, const uint64_t* __restrict__ ranges) {
    __shared__ uint32_t s_rk[SM_RCAP];
    __shared__ uint32_t s_x[SM_TILE + SM_TILE / 16];   // probe keys, then build row (or SM_MISS) per probe row
    __shared__ uint32_t s_w[SM_WAVES];
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t tile = blockIdx.x;
    const uint64_t base = tile * SM_TILE;
    const uint32_t cnt_rows = (uint32_t)min((uint64_t)SM_TILE, ns - base);
    // probe rows base + i*BLOCK + t (coalesced); keys to LDS, payloads kept
    uint32_t pay[SM_ITEMS];
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const uint32_t r = (uint32_t)i * SM_BLOCK + t;
        const uint32_t k = r < cnt_rows ? sk[base + r] : 0xFFFFFFFFu;
        pay[i] = r < cnt_rows ? sp[base + r] : 0u;
        s_x[smj_pad(r)] = k;
    }
    const uint64_t lo = ranges[2 * tile], hi = ranges[2 * tile + 1];
    const uint32_t m = (uint32_t)(hi - lo);
    const bool cached = m <= SM_RCAP;
    if (cached) {
        // every load of the build range first, then the LDS stores
        uint32_t bk[SM_RCAP / SM_BLOCK];
#pragma unroll
        for (int k = 0; k < (int)(SM_RCAP / SM_BLOCK); k++) {
            const uint32_t i = (uint32_t)k * SM_BLOCK + t;
            bk[k] = i < m ? rk[lo + i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < (int)(SM_RCAP / SM_BLOCK); k++) {
            const uint32_t i = (uint32_t)k * SM_BLOCK + t;
            if (i < m) s_rk[i] = bk[k];
        }
    }
    __syncthreads();
    auto bkey = [&](uint32_t p) -> uint32_t { return cached ? s_rk[p] : rk[lo + p]; };
    // ---- merge: rows t*16 .. t*16+15 of the tile ----
    {
        const uint32_t r0 = t * SM_ITEMS;
        uint32_t p = 0;
        if (r0 < cnt_rows) {
            const uint32_t k0 = s_x[smj_pad(r0)];
            uint32_t a = 0, b = m;   // lower bound of k0 in the build range
            while (a < b) {
                const uint32_t mid = (a + b) >> 1;
                if (bkey(mid) < k0) a = mid + 1; else b = mid;
            }
            p = a;
        }
#pragma unroll 4
        for (int j = 0; j < SM_ITEMS; j++) {
            const uint32_t r = r0 + j;
            if (r >= cnt_rows) break;
            const uint32_t k = s_x[smj_pad(r)];
            while (p < m && bkey(p) < k) p++;
            s_x[smj_pad(r)] = (p < m && bkey(p) == k) ? (uint32_t)(lo + p) : SM_MISS;
        }
    }
    __syncthreads();
    // ---- gather build payloads (all in flight together), then write the
    // matched pairs at their rows ----
    uint32_t bval[SM_ITEMS], hits = 0;
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const uint32_t r = (uint32_t)i * SM_BLOCK + t;
        const uint32_t bx = r < cnt_rows ? s_x[smj_pad(r)] : SM_MISS;
        hits |= (uint32_t)(bx != SM_MISS) << i;
        bval[i] = bx != SM_MISS ? rp[bx] : 0u;
    }
    uint32_t matched = 0;
#pragma unroll
    for (int i = 0; i < SM_ITEMS; i++) {
        const uint32_t r = (uint32_t)i * SM_BLOCK + t;
        const bool hit = (hits >> i) & 1;
        if (hit) {
            out_p[base + r] = pay[i];
            out_b[base + r] = bval[i];
            matched++;
        }
        const uint64_t bal = pd::ballot(hit);
        if (l == 0 && base + (uint64_t)i * SM_BLOCK + (uint64_t)w * PD_WAVE < ns)
            valid[(base + (uint64_t)i * SM_BLOCK + (uint64_t)w * PD_WAVE) / PD_WAVE] = bal;
    }
    matched = pd::wave_sum(matched);
    if (l == 0) s_w[w] = matched;
    __syncthreads();
    if (t == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < SM_WAVES; i++) tot += s_w[i];
        tcnt[tile] = tot;
    }
}

// Packs the direct merge's output in probe order: tile `tile`'s matched rows
// (bitmap words valid[tile * 64 ..]) go to toff[tile] onwards.
__global__ __launch_bounds__(SM_BLOCK) void k_smj_pack(uint64_t ns, const uint64_t* __restrict__ valid,
                                                       const uint32_t* __restrict__ toff,
                                                       const uint32_t* __restrict__ in_p, const uint32_t* __restrict__ in_b,
                                                       uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                       uint64_t cap) {
    constexpr uint32_t WPT = SM_TILE / PD_WAVE;   // 64 bitmap words per tile
    __shared__ uint32_t s_wo[WPT];
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t tile = blockIdx.x;
    const uint64_t w0 = tile * WPT;
    const uint64_t nw_all = (ns + PD_WAVE - 1) / PD_WAVE;
    if (w == 0) {
        const uint32_t c = (w0 + l < nw_all) ? pd::popc64(valid[w0 + l]) : 0u;
        s_wo[l] = pd::wave_incl_scan(c) - c;
    }
    __syncthreads();
    const uint64_t obase = toff[tile];
    const uint64_t lt = pd::lanemask_lt();
    for (uint32_t j = w; j < WPT && w0 + j < nw_all; j += SM_WAVES) {
        const uint64_t bits = valid[w0 + j];
        if ((bits >> l) & 1) {
            const uint64_t src = (w0 + j) * PD_WAVE + l;
            const uint64_t dst = obase + s_wo[j] + pd::popc64(bits & lt);
            if (dst < cap) {
                out_p[dst] = in_p[src];
                out_b[dst] = in_b[src];
            }
        }
    }
}

// Build-side range each probe tile can match: [lower_bound(first key),
// upper_bound(last key)).  One thread per tile, all tiles in parallel.
__global__ void k_smj_tile_ranges(const uint32_t* __restrict__ rk, uint64_t nr, const uint32_t* __restrict__ sk,
                                  uint64_t ns, uint64_t ntiles, uint64_t* __restrict__ ranges) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = t * SM_TILE, e = min(ns, b + SM_TILE);
        const uint64_t lo = lower_bound_u32(rk, 0, nr, sk[b]);
        ranges[2 * t] = lo;
        ranges[2 * t + 1] = upper_bound_u32(rk, lo, nr, sk[e - 1]);
    }
}

// per probe row: first matching build row and number of matches
__global__ void k_smj_ranges(const uint32_t* __restrict__ rk, uint64_t nr, const uint32_t* __restrict__ sk, uint64_t ns,
                             uint32_t* __restrict__ first, uint32_t* __restrict__ cnt) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ns; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t key = sk[j];
        const uint64_t a = lower_bound_u32(rk, 0, nr, key);
        const uint64_t b = upper_bound_u32(rk, a, nr, key);
        first[j] = (uint32_t)a;
        cnt[j] = (uint32_t)(b - a);
    }
}

__global__ void k_smj_expand(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ sp, uint64_t ns,
                             const uint32_t* __restrict__ first, const uint32_t* __restrict__ cnt,
                             const uint32_t* __restrict__ off, uint32_t* __restrict__ out_p,
                             uint32_t* __restrict__ out_b, uint64_t cap) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ns; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = cnt[j];
        const uint64_t o = off[j];
        const uint32_t f = first[j];
        const uint32_t p = sp[j];
        for (uint32_t e = 0; e < c; e++) {
            if (o + e < cap) {
                out_p[o + e] = p;
                out_b[o + e] = rp[f + e];
            }
        }
    }
}

uint64_t sort_merge_join(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                         const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap,
                         hipStream_t s) {
    PD_REQUIRE(nb < (1ull << 30) && np < (1ull << 30), PIMDAL_E_INVALID, "sort-merge join inputs must have < 2^30 rows");
    PD_REQUIRE((bk || !nb) && (pk || !np), PIMDAL_E_INVALID, "key pointer is NULL");
    if (nb == 0 || np == 0) return 0;
    Scratch rk(c, nb * 4, s), rp(c, nb * 4, s), sk(c, np * 4, s), spay(c, np * 4, s);
    Scratch rid_b(c, bp ? 0 : nb * 4, s), rid_p(c, pp ? 0 : np * 4, s);
    // a missing payload means "the key itself", as in the reference's keys-only join
    sort_pairs(c, bk, bp ? (const void*)bp : (const void*)bk, 4, nb, rk.as<uint32_t>(), rp.get(), -1, -1, s);
    sort_pairs(c, pk, pp ? (const void*)pp : (const void*)pk, 4, np, sk.as<uint32_t>(), spay.get(), -1, -1, s);
    Scratch small(c, 16, s);
    fill_u32(c, small.get(), 0, 4, s);
    const uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(nb, 256), (uint64_t)c->cu_count * 8);
    launch(c, "smj_dupcheck", k_adjacent_dup, dim3(std::max(1u, g)), dim3(256), 0, s, (const uint32_t*)rk.as<uint32_t>(),
           nb, small.as<uint32_t>());
    uint32_t dup = 0;
    read_small(c, &dup, small.get(), 4, s);
    if (!dup) {
        const uint64_t ntiles = ceil_div(np, SM_TILE);
        Scratch st(c, 8 * (ntiles + 1), s);
        fill_u64(c, st.get(), 0, ntiles + 1, s);
        const bool direct = cap >= np && out_p && out_b;
        Scratch tp(c, direct ? 0 : np * 4, s), tb(c, direct ? 0 : np * 4, s);
        uint32_t* op = direct ? out_p : tp.as<uint32_t>();
        uint32_t* ob = direct ? out_b : tb.as<uint32_t>();
        Scratch rng(c, ntiles * 16, s);
        launch(c, "smj_ranges", k_smj_tile_ranges, dim3((uint32_t)ceil_div(ntiles, 256)), dim3(256), 0, s,
               (const uint32_t*)rk.as<uint32_t>(), nb, (const uint32_t*)sk.as<uint32_t>(), np, ntiles,
               rng.as<uint64_t>());
        // merge (knob smj_merge): 1 = direct output slots + match bitmap,
        // 0 = compacting merge with decoupled look-back
        if (c->tune_get("smj_merge", 1) == 1) {
            Scratch valid(c, ceil_div(np, 64) * 8, s), tcnt(c, (ntiles + 1) * 4, s), toff(c, (ntiles + 1) * 4, s),
                tot(c, 8, s);
            launch(c, "smj_merge", k_smj_direct, dim3((uint32_t)ntiles), dim3(SM_BLOCK), 0, s,
                   (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(),
                   (const uint32_t*)sk.as<uint32_t>(), (const uint32_t*)spay.as<uint32_t>(), np, op, ob,
                   valid.as<uint64_t>(), tcnt.as<uint32_t>(), (const uint64_t*)rng.as<uint64_t>());
            exclusive_scan_u32(c, tcnt.as<uint32_t>(), toff.as<uint32_t>(), ntiles, tot.as<uint64_t>(), s);
            const uint64_t total = read_u64(c, tot.get(), s);
            if (total == np) {   // every probe row matched: complete and in order as written
                if (!direct && out_p && out_b && cap) {
                    copy_d2d(c, out_p, op, std::min(total, cap) * 4, s);
                    copy_d2d(c, out_b, ob, std::min(total, cap) * 4, s);
                    PD_HIP(hipStreamSynchronize(s));
                }
                return total;
            }
            if (out_p && out_b && cap) {
                Scratch cp(c, direct ? total * 4 + 4 : 0, s), cb(c, direct ? total * 4 + 4 : 0, s);
                launch(c, "smj_pack", k_smj_pack, dim3((uint32_t)ntiles), dim3(SM_BLOCK), 0, s, np,
                       (const uint64_t*)valid.as<uint64_t>(), (const uint32_t*)toff.as<uint32_t>(), (const uint32_t*)op,
                       (const uint32_t*)ob, direct ? cp.as<uint32_t>() : out_p, direct ? cb.as<uint32_t>() : out_b,
                       direct ? total : cap);
                if (direct) {
                    copy_d2d(c, out_p, cp.get(), std::min(total, cap) * 4, s);
                    copy_d2d(c, out_b, cb.get(), std::min(total, cap) * 4, s);
                }
                PD_HIP(hipStreamSynchronize(s));
            }
            return total;
        }
        launch(c, "smj_merge", k_smj_unique, dim3((uint32_t)ntiles), dim3(SM_BLOCK), 0, s,
               (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(), nb,
               (const uint32_t*)sk.as<uint32_t>(), (const uint32_t*)spay.as<uint32_t>(), np, op, ob,
               st.as<uint64_t>(), reinterpret_cast<uint32_t*>(st.as<uint64_t>() + ntiles), ntiles,
               small.as<uint64_t>() + 1, (const uint64_t*)rng.as<uint64_t>());
        uint64_t total = 0;
        read_small(c, &total, small.as<uint64_t>() + 1, 8, s);
        if (!direct && out_p && out_b && cap) {
            copy_d2d(c, out_p, op, std::min(total, cap) * 4, s);
            copy_d2d(c, out_b, ob, std::min(total, cap) * 4, s);
            PD_HIP(hipStreamSynchronize(s));
        }
        return total;
    }
    Scratch first(c, np * 4, s), cnt(c, (np + 1) * 4, s), off(c, (np + 1) * 4, s), tot(c, 8, s);
    const uint32_t gp = (uint32_t)std::min<uint64_t>(ceil_div(np, 256), (uint64_t)c->cu_count * 16);
    launch(c, "smj_ranges", k_smj_ranges, dim3(gp), dim3(256), 0, s, (const uint32_t*)rk.as<uint32_t>(), nb,
           (const uint32_t*)sk.as<uint32_t>(), np, first.as<uint32_t>(), cnt.as<uint32_t>());
    exclusive_scan_u32(c, cnt.as<uint32_t>(), off.as<uint32_t>(), np, tot.as<uint64_t>(), s);
    uint64_t total = read_u64(c, tot.get(), s);
    PD_REQUIRE(total < (1ull << 32), PIMDAL_E_UNSUPPORTED, "join output exceeds 2^32-1 pairs");
    if (out_p && out_b && cap)
        launch(c, "smj_expand", k_smj_expand, dim3(gp), dim3(256), 0, s, (const uint32_t*)rp.as<uint32_t>(),
               (const uint32_t*)spay.as<uint32_t>(), np, (const uint32_t*)first.as<uint32_t>(),
               (const uint32_t*)cnt.as<uint32_t>(), (const uint32_t*)off.as<uint32_t>(), out_p, out_b, cap);
    PD_HIP(hipStreamSynchronize(s));
    return total;
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_sort_merge_join_dev(pimdal_ctx* c, const uint32_t* build_key, const uint32_t* build_payload,
                                          uint64_t n_build, const uint32_t* probe_key, const uint32_t* probe_payload,
                                          uint64_t n_probe, uint32_t* out_probe_payload, uint32_t* out_build_payload,
                                          uint64_t out_capacity, uint64_t* out_count) {
    PD_API_BEGIN
    PD_REQUIRE(c && out_count, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    const uint64_t k = sort_merge_join(c, build_key, build_payload, n_build, probe_key, probe_payload, n_probe,
                                       out_probe_payload, out_build_payload, out_capacity, c->stream);
    *out_count = k;
    PD_REQUIRE(k <= out_capacity, PIMDAL_E_CAPACITY, "join output larger than out_capacity (count returned)");
    PD_API_END
}

// ---------------- column-level joins (library allocates the outputs) ----------------

namespace {


int join_cols(JoinFn fn, pimdal_ctx* c, const pimdal_col* bk, const pimdal_col* bp, const pimdal_col* pk,
              const pimdal_col* pp, pimdal_col** op, pimdal_col** ob) {
    PD_API_BEGIN
    PD_REQUIRE(c && bk && pk && op && ob, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    PD_REQUIRE(type_width(bk->type) == 4 && type_width(pk->type) == 4, PIMDAL_E_UNSUPPORTED, "join keys must be 32-bit");
    PD_REQUIRE(!bp || (bp->length == bk->length && type_width(bp->type) == 4), PIMDAL_E_INVALID, "bad build payload");
    PD_REQUIRE(!pp || (pp->length == pk->length && type_width(pp->type) == 4), PIMDAL_E_INVALID, "bad probe payload");
    if (bk->validity || pk->validity || (bp && bp->validity) || (pp && pp->validity)) {
        // NULL keys join nothing; payload validity follows the pairs
        join_cols_nullable(fn, c, bk, bp, pk, pp, op, ob);
        return PIMDAL_OK;
    }
    // first call sizes the output (probe-sized buffers cover the unique-build case in one pass)
    uint64_t cap = pk->length;
    pimdal_col* a = nullptr;
    pimdal_col* b = nullptr;
    auto alloc = [&](uint64_t n) {
        a = new pimdal_col{pp ? pp->type : pk->type, 0, n, c->pool.alloc(std::max<uint64_t>(n, 1) * 4, c->stream), true};
        b = new pimdal_col{bp ? bp->type : bk->type, 0, n, c->pool.alloc(std::max<uint64_t>(n, 1) * 4, c->stream), true};
    };
    auto release = [&]() {
        if (a) { c->pool.free(a->data, c->stream); delete a; a = nullptr; }
        if (b) { c->pool.free(b->data, c->stream); delete b; b = nullptr; }
    };
    try {
        alloc(cap);
        uint64_t k = fn(c, (const uint32_t*)bk->data, bp ? (const uint32_t*)bp->data : nullptr, bk->length,
                        (const uint32_t*)pk->data, pp ? (const uint32_t*)pp->data : nullptr, pk->length,
                        (uint32_t*)a->data, (uint32_t*)b->data, cap, c->stream);
        if (k > cap) {
            release();
            alloc(k);
            k = fn(c, (const uint32_t*)bk->data, bp ? (const uint32_t*)bp->data : nullptr, bk->length,
                   (const uint32_t*)pk->data, pp ? (const uint32_t*)pp->data : nullptr, pk->length,
                   (uint32_t*)a->data, (uint32_t*)b->data, k, c->stream);
        }
        PD_HIP(hipStreamSynchronize(c->stream));
        a->length = b->length = k;
    } catch (...) {
        release();
        throw;
    }
    *op = a;
    *ob = b;
    PD_API_END
}

}  // namespace

extern "C" int pimdal_hash_join(pimdal_ctx* c, const pimdal_col* bk, const pimdal_col* bp, const pimdal_col* pk,
                                const pimdal_col* pp, pimdal_col** op, pimdal_col** ob) {
    return join_cols(pd::hash_join, c, bk, bp, pk, pp, op, ob);
}

extern "C" int pimdal_sort_merge_join(pimdal_ctx* c, const pimdal_col* bk, const pimdal_col* bp, const pimdal_col* pk,
                                      const pimdal_col* pp, pimdal_col** op, pimdal_col** ob) {
    return join_cols(pd::sort_merge_join, c, bk, bp, pk, pp, op, ob);
}
