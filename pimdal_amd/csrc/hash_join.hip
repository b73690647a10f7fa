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
// Plan on gfx950 (up to 2^24 build rows, knob hj_onepass):
//  1. one partition pass per side (partition.hip) into F <= 1024 partitions
//     of about 16384 build rows by the reference hash: per-tile counts, a
//     scan over tiles, one scatter of (key, payload) rows in partition runs;
//  2. work items of whole probe partitions (about 131072 rows; hot ones
//     split);
//  3. per item (k_hj_csr) the build partition bucketed in LDS by a counting
//     sort (8192 buckets), every probe row's pair written at its own
//     position with a match bitmap (packed afterwards unless every row
//     matched).  Duplicate build keys or a partition past the table fall
//     back to the general plan.
// General plan:
//  1. radix-partition both relations by the reference hash
//     (hash_join.cpp:20-25) into F partitions with the onesweep passes of
//     the radix sort run over hash digits (hash_partition): each pass
//     reorders an 8192-row tile by digit in LDS and writes digit runs, so
//     HBM sees long sequential runs instead of one 8-byte write per row
//     into F x blocks separate streams; partition bounds come from a
//     binary search over the partitioned keys;
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
#include <vector>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int HJ_BLOCK = 512;         // count/write phases; the unique phase takes its shape as parameters
constexpr int HJ_TS = 8192;           // LDS table slots (8 B each)
constexpr int HJ_ROWS = 16;           // probe rows per thread per round

// ---------------- work items over probe partitions ----------------

// The probe's plan in one block: wst[f] = first work item of probe
// partition f (ceil(rows / CH) items each, an exclusive scan, wst[F] =
// total), and the flags and per-item counts cleared.  F <= 16384.
__global__ __launch_bounds__(1024) void k_hj_plan(const uint32_t* __restrict__ HS, uint32_t F, uint32_t CH,
                                                  uint32_t* __restrict__ wst, uint32_t* __restrict__ flags,
                                                  uint32_t* __restrict__ wcnt, uint32_t nwcnt) {
    __shared__ uint32_t s_w[1024 / PD_WAVE];
    const uint32_t t = threadIdx.x, w = t / PD_WAVE;
    for (uint32_t i = t; i < nwcnt; i += 1024) wcnt[i] = 0;
    if (t < 4) flags[t] = 0;
    const uint32_t per = (F + 1023) / 1024, f0 = min(F, t * per), f1 = min(F, f0 + per);
    auto items = [&](uint32_t f) -> uint32_t { return (HS[f + 1] - HS[f] + CH - 1) / CH; };
    uint32_t sum = 0;
    for (uint32_t f = f0; f < f1; f++) sum += items(f);
    const uint32_t inc = pd::wave_incl_scan(sum);
    if (pd::lane() == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t base = inc - sum;
    for (uint32_t i = 0; i < w; i++) base += s_w[i];
    for (uint32_t f = f0; f < f1; f++) {
        wst[f] = base;
        base += items(f);
    }
    if (t == 1023) wst[F] = base;   // the last thread's range ends at F
}

// flags[0]: a probe row matched more than one build row
// flags[1]: a build partition exceeded the LDS table
// flags[2..3]: total emitted pairs (uint64)
enum { HJ_UNIQUE = 0, HJ_COUNT = 1, HJ_WRITE = 2 };

// BLOCK threads x ROWS probe rows per thread per round (BLOCK / 64 * ROWS == 128).
template <int PHASE, int BLOCK = HJ_BLOCK, int ROWS = HJ_ROWS>
__global__ __launch_bounds__(BLOCK, BLOCK == 1024 ? 8 : 2) void k_hj_join(const uint32_t* __restrict__ Rk,
                                                         const uint32_t* __restrict__ Rp, const uint32_t* __restrict__ HR,
                                                         uint32_t GR, const uint32_t* __restrict__ Sk,
                                                         const uint32_t* __restrict__ Sp,
                                                         const uint32_t* __restrict__ HS, uint32_t GS, uint32_t F,
                                                         uint32_t CH, const uint32_t* __restrict__ wstart,
                                                         uint32_t* __restrict__ wcount,
                                                         const uint32_t* __restrict__ woff,
                                                         uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                         uint64_t out_cap, uint32_t* __restrict__ flags) {
    __shared__ uint2 tab[HJ_TS];
    __shared__ uint32_t occ[HJ_TS / 32];
    __shared__ uint32_t s_wsum[BLOCK / PD_WAVE];
    __shared__ uint32_t s_kc[ROWS * (BLOCK / PD_WAVE)];
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
    for (uint32_t i = t; i < HJ_TS / 32; i += BLOCK) occ[i] = 0;
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
    for (uint64_t i = r0 + t; i < r1; i += BLOCK) {
        const uint2 kv = make_uint2(Rk[i], Rp[i]);
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

    // ---- probe in rounds of BLOCK * ROWS rows ----
    uint64_t written = 0;        // block-uniform running output count
    const uint64_t wbase = (PHASE == HJ_WRITE) ? woff[wi] : s0;
    bool dup = false;
    for (uint64_t rb = s0; rb < s1; rb += (uint64_t)BLOCK * ROWS) {
        uint32_t key[ROWS], pay[ROWS], first[ROWS], m[ROWS];
        uint32_t cnt = 0;
        // all loads of the round first, so they are in flight together
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            const uint64_t i = rb + (uint64_t)k * BLOCK + t;
            key[k] = i < s1 ? Sk[i] : 0u;
            pay[k] = i < s1 ? Sp[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            const uint64_t i = rb + (uint64_t)k * BLOCK + t;
            m[k] = 0;
            first[k] = 0;
            if (i < s1) {
                uint32_t h = pd::slot_hash(key[k]) & (HJ_TS - 1);
                while (occ[h >> 5] & (1u << (h & 31))) {
                    const uint2 e = tab[h];
                    if (e.x == key[k]) {
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
        if constexpr (PHASE == HJ_UNIQUE) {
            // at most one match per row: output slots in row order
            // (row = rb + k * BLOCK + wave * 64 + lane), so every store
            // instruction writes consecutive slots
            constexpr int WAVES = BLOCK / PD_WAVE;
            static_assert(ROWS * WAVES == 2 * PD_WAVE, "one wave scans the (row group, wave) counts");
#pragma unroll
            for (int k = 0; k < ROWS; k++) {
                const uint64_t b = pd::ballot(m[k] != 0);
                if (l == 0) s_kc[k * WAVES + w] = pd::popc64(b);
            }
            __syncthreads();
            if (w == 0) {
                const uint32_t a0 = s_kc[2 * l], a1 = s_kc[2 * l + 1];
                const uint32_t inc = pd::wave_incl_scan(a0 + a1);
                s_kc[2 * l] = inc - a0 - a1;
                s_kc[2 * l + 1] = inc - a1;
                if (l == PD_WAVE - 1) s_wsum[0] = inc;
            }
            __syncthreads();
            const uint32_t tot = s_wsum[0];
            const uint64_t lt = pd::lanemask_lt();
#pragma unroll
            for (int k = 0; k < ROWS; k++) {
                const uint64_t b = pd::ballot(m[k] != 0);
                if (m[k]) {
                    const uint64_t o = wbase + written + s_kc[k * WAVES + w] + pd::popc64(b & lt);
                    if (o < out_cap) { out_p[o] = pay[k]; out_b[o] = first[k]; }
                }
            }
            written += tot;
            __syncthreads();   // s_kc / s_wsum are rewritten by the next round
            continue;
        }
        // block exclusive scan of cnt
        const uint32_t inc = pd::wave_incl_scan(cnt);
        if (l == PD_WAVE - 1) s_wsum[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < BLOCK / PD_WAVE; i++) {
            const uint32_t v = s_wsum[i];
            pre += (i < (int)w) ? v : 0;
            tot += v;
        }
        __syncthreads();
        if (PHASE != HJ_COUNT) {
            uint64_t o = wbase + written + pre + inc - cnt;
#pragma unroll
            for (int k = 0; k < ROWS; k++) {
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

// Unique-key probe with direct output slots: a matching probe row at
// partitioned position i writes its pair to slot i, and a bitmap records
// which rows matched (item-local: word j of item wi covers rows s0 + 64 j ..
// s0 + 64 j + 63, at valid[wi * vw + j]).  With every probe row matched (a
// foreign-key join) the output is complete and dense as written; otherwise
// k_hj_compact_bits packs it.  No block-wide scan per round (the compacting
// probe above needs three barriers per round), and each round's keys and
// payloads are loaded while the previous round probes.  flags as in
// k_hj_join; wcount[wi] = rows of the item that matched.
template <int BLOCK, int ROWS>
__global__ __launch_bounds__(BLOCK, BLOCK == 1024 ? 8 : 2) void k_hj_direct(
    const uint32_t* __restrict__ Rk, const uint32_t* __restrict__ Rp, const uint32_t* __restrict__ HR,
    const uint32_t* __restrict__ Sk, const uint32_t* __restrict__ Sp, const uint32_t* __restrict__ HS, uint32_t F,
    uint32_t CH, const uint32_t* __restrict__ wstart, uint32_t* __restrict__ wcount, uint32_t* __restrict__ out_p,
    uint32_t* __restrict__ out_b, uint64_t* __restrict__ valid, uint32_t vw, uint32_t* __restrict__ flags) {
    constexpr int WAVES = BLOCK / PD_WAVE;
    __shared__ uint2 tab[HJ_TS];
    __shared__ uint32_t occ[HJ_TS / 32];
    __shared__ uint32_t s_f, s_dup, s_cnt[WAVES];

    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (t == 0) {
        uint32_t lo = 0, hi = F;   // partition: last f with wstart[f] <= wi
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_f = lo;
        s_dup = 0;
    }
    for (uint32_t i = t; i < HJ_TS / 32; i += BLOCK) occ[i] = 0;
    __syncthreads();
    const uint32_t f = s_f;
    const uint32_t ck = wi - wstart[f];
    const uint64_t r0 = HR[f], r1 = HR[f + 1];
    const uint64_t s0 = (uint64_t)HS[f] + (uint64_t)ck * CH, s1 = min((uint64_t)HS[f + 1], s0 + CH);
    if (r1 - r0 > (uint64_t)HJ_TS * 7 / 8) {
        if (t == 0) {
            atomicOr(&flags[1], 1u);
            wcount[wi] = 0;
        }
        return;
    }
    // first round's probe rows, in flight during the build
    uint32_t key[ROWS], pay[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
        const uint64_t i = s0 + (uint64_t)k * BLOCK + t;
        key[k] = i < s1 ? Sk[i] : 0u;
        pay[k] = i < s1 ? Sp[i] : 0u;
    }
    // ---- build: every build row of this thread loaded first, then inserted
    // (a partition holds at most 7/8 of the table, so BR rows per thread) ----
    constexpr int BR = (HJ_TS * 7 / 8 + BLOCK - 1) / BLOCK;
    uint2 brow[BR];
#pragma unroll
    for (int k = 0; k < BR; k++) {
        const uint64_t i = r0 + (uint64_t)k * BLOCK + t;
        brow[k] = i < r1 ? make_uint2(Rk[i], Rp[i]) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < BR; k++) {
        if (r0 + (uint64_t)k * BLOCK + t >= r1) break;
        uint32_t h = pd::slot_hash(brow[k].x) & (HJ_TS - 1);
        while (true) {
            const uint32_t bit = 1u << (h & 31);
            const uint32_t prev = atomicOr(&occ[h >> 5], bit);
            if (!(prev & bit)) {
                tab[h] = brow[k];
                break;
            }
            h = (h + 1) & (HJ_TS - 1);
        }
    }
    __syncthreads();

    // ---- probe ----
    uint32_t matched = 0;
    bool dup = false;
    for (uint64_t rb = s0; rb < s1; rb += (uint64_t)BLOCK * ROWS) {
        uint32_t nkey[ROWS], npay[ROWS];
        const uint64_t nb = rb + (uint64_t)BLOCK * ROWS;
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            const uint64_t i = nb + (uint64_t)k * BLOCK + t;
            nkey[k] = i < s1 ? Sk[i] : 0u;
            npay[k] = i < s1 ? Sp[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            const uint64_t i = rb + (uint64_t)k * BLOCK + t;
            uint32_t m = 0, first = 0;
            if (i < s1) {
                uint32_t h = pd::slot_hash(key[k]) & (HJ_TS - 1);
                uint32_t o = occ[h >> 5];
                uint2 e = tab[h];
                while (o & (1u << (h & 31))) {
                    if (e.x == key[k]) {
                        if (m == 0) first = e.y;
                        m++;
                    }
                    h = (h + 1) & (HJ_TS - 1);
                    o = occ[h >> 5];
                    e = tab[h];
                }
            }
            dup |= m > 1;
            if (m) {
                out_p[i] = pay[k];
                out_b[i] = first;
                matched++;
            }
            // rows rb - s0 + k * BLOCK + w * 64 .. + 63 of the item: one word
            const uint64_t b = pd::ballot(m != 0);
            if (l == 0) valid[(uint64_t)wi * vw + (rb - s0 + (uint64_t)k * BLOCK + (uint64_t)w * PD_WAVE) / PD_WAVE] = b;
        }
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            key[k] = nkey[k];
            pay[k] = npay[k];
        }
    }
    matched = pd::wave_sum(matched);
    if (l == 0) s_cnt[w] = matched;
    if (dup) s_dup = 1;
    __syncthreads();
    if (t == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < WAVES; i++) tot += s_cnt[i];
        wcount[wi] = tot;
        if (s_dup) atomicOr(&flags[0], 1u);
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(flags + 2), (unsigned long long)tot);
    }
}

// The direct probe for partitions of up to HC_RCAP build rows (one-pass
// partitioning gives ~16384 rows per partition): the partition's rows are
// bucketed in LDS by a counting sort (8192 buckets, 16-bit counts packed two
// to a word, ranks from the atomic adds, an exclusive scan into bucket
// starts), so the table holds exactly its rows and a probe walks one
// bucket, about two rows.  Output, bitmap, counts and flags as k_hj_direct.
constexpr uint32_t HC_RCAP = 16896;
constexpr uint32_t HC_NB = 8192;
constexpr int HC_BLOCK = 1024;

__device__ __forceinline__ uint32_t hc_bucket(uint32_t k) { return pd::slot_hash(k) & (HC_NB - 1); }

// HC_ROWS probe rows per thread per round (the next round's rows load
// while this round probes).
template <int HC_ROWS>
__global__ __launch_bounds__(HC_BLOCK) void k_hj_csr(
    const uint2* __restrict__ R, const uint32_t* __restrict__ HR, const uint2* __restrict__ S,
    const uint32_t* __restrict__ HS, uint32_t F,
    uint32_t CH, const uint32_t* __restrict__ wstart, uint32_t* __restrict__ wcount, uint32_t* __restrict__ out_p,
    uint32_t* __restrict__ out_b, uint64_t* __restrict__ valid, uint32_t vw, uint32_t* __restrict__ flags) {
    constexpr int WAVES = HC_BLOCK / PD_WAVE;
    constexpr int BR = (HC_RCAP + HC_BLOCK - 1) / HC_BLOCK;   // build rows per thread
    __shared__ uint2 tab[HC_RCAP];
    __shared__ uint32_t bw[HC_NB / 2 + 1];   // 16-bit bucket counts, then starts; start[HC_NB] = rows
    __shared__ uint32_t s_f, s_dup, s_cnt[WAVES];
    const uint16_t* bst = reinterpret_cast<const uint16_t*>(bw);

    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (t == 0) {
        uint32_t lo = 0, hi = F;   // partition: last f with wstart[f] <= wi
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_f = lo;
        s_dup = 0;
    }
    for (uint32_t i = t; i < HC_NB / 2 + 1; i += HC_BLOCK) bw[i] = 0;
    __syncthreads();
    const uint32_t f = s_f;
    const uint32_t ck = wi - wstart[f];
    const uint64_t r0 = HR[f], r1 = HR[f + 1];
    const uint64_t s0 = (uint64_t)HS[f] + (uint64_t)ck * CH, s1 = min((uint64_t)HS[f + 1], s0 + CH);
    if (r1 - r0 > HC_RCAP) {
        if (t == 0) {
            atomicOr(&flags[1], 1u);
            wcount[wi] = 0;
        }
        return;
    }
    // first round's probe rows, in flight during the build
    uint2 row[HC_ROWS];
#pragma unroll
    for (int k = 0; k < HC_ROWS; k++) {
        const uint64_t i = s0 + (uint64_t)k * HC_BLOCK + t;
        row[k] = i < s1 ? S[i] : make_uint2(0u, 0u);
    }
    // ---- build: count per bucket (rank = the count before this row) ----
    uint2 brow[BR];
    uint32_t brank[BR];
#pragma unroll
    for (int k = 0; k < BR; k++) {
        const uint64_t i = r0 + (uint64_t)k * HC_BLOCK + t;
        brow[k] = i < r1 ? R[i] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < BR; k++) {
        if (r0 + (uint64_t)k * HC_BLOCK + t >= r1) break;
        const uint32_t b = hc_bucket(brow[k].x), sh = (b & 1) << 4;
        brank[k] = (atomicAdd(&bw[b >> 1], 1u << sh) >> sh) & 0xFFFFu;
    }
    __syncthreads();
    // bucket starts: thread t scans buckets 8t .. 8t + 7 (words 4t .. 4t + 3)
    {
        uint32_t cnt[8], sum = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t x = bw[4 * t + q];
            cnt[2 * q] = x & 0xFFFFu;
            cnt[2 * q + 1] = x >> 16;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) sum += cnt[q];
        const uint32_t inc = pd::wave_incl_scan(sum);
        if (l == PD_WAVE - 1) s_cnt[w] = inc;
        __syncthreads();
        uint32_t run = inc - sum;
        for (int i = 0; i < (int)w; i++) run += s_cnt[i];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t a = run, b = run + cnt[2 * q];
            run = b + cnt[2 * q + 1];
            bw[4 * t + q] = a | (b << 16);
        }
        if (t == HC_BLOCK - 1) bw[HC_NB / 2] = run;   // start[HC_NB] = rows
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BR; k++) {
        if (r0 + (uint64_t)k * HC_BLOCK + t >= r1) break;
        tab[bst[hc_bucket(brow[k].x)] + brank[k]] = brow[k];
    }
    __syncthreads();

    // ---- probe ----
    uint32_t matched = 0;
    bool dup = false;
    for (uint64_t rb = s0; rb < s1; rb += (uint64_t)HC_BLOCK * HC_ROWS) {
        uint2 nrow[HC_ROWS];
        const uint64_t nb = rb + (uint64_t)HC_BLOCK * HC_ROWS;
#pragma unroll
        for (int k = 0; k < HC_ROWS; k++) {
            const uint64_t i = nb + (uint64_t)k * HC_BLOCK + t;
            nrow[k] = i < s1 ? S[i] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < HC_ROWS; k++) {
            const uint64_t i = rb + (uint64_t)k * HC_BLOCK + t;
            uint32_t m = 0, first = 0;
            if (i < s1) {
                const uint32_t b = hc_bucket(row[k].x);
                const uint32_t e = bst[b + 1];
                for (uint32_t j = bst[b]; j < e; j++) {
                    const uint2 r = tab[j];
                    if (r.x == row[k].x) {
                        if (m == 0) first = r.y;
                        m++;
                    }
                }
            }
            dup |= m > 1;
            if (m) {
                out_p[i] = row[k].y;
                out_b[i] = first;
                matched++;
            }
            const uint64_t bb = pd::ballot(m != 0);
            if (l == 0) valid[(uint64_t)wi * vw + (rb - s0 + (uint64_t)k * HC_BLOCK + (uint64_t)w * PD_WAVE) / PD_WAVE] = bb;
        }
#pragma unroll
        for (int k = 0; k < HC_ROWS; k++) row[k] = nrow[k];
    }
    matched = pd::wave_sum(matched);
    if (l == 0) s_cnt[w] = matched;
    if (dup) s_dup = 1;
    __syncthreads();
    if (t == 0) {
        uint32_t tot = 0;
        for (int i = 0; i < WAVES; i++) tot += s_cnt[i];
        wcount[wi] = tot;
        if (s_dup) atomicOr(&flags[0], 1u);
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(flags + 2), (unsigned long long)tot);
    }
}

// Packs the direct probe's output: work item wi's matched rows (bitmap
// words valid[wi * vw ..]) go, in row order, to woff[wi] onwards.  One block
// per work item; wave v takes words v, v + 4, ...; lane = bit.
__global__ __launch_bounds__(256) void k_hj_compact_bits(const uint32_t* __restrict__ HS, uint32_t F, uint32_t CH,
                                                         const uint32_t* __restrict__ wstart,
                                                         const uint32_t* __restrict__ woff,
                                                         const uint64_t* __restrict__ valid, uint32_t vw,
                                                         const uint32_t* __restrict__ in_p,
                                                         const uint32_t* __restrict__ in_b, uint32_t* __restrict__ out_p,
                                                         uint32_t* __restrict__ out_b, uint64_t out_cap) {
    __shared__ uint32_t s_f, s_w[4];
    __shared__ uint32_t s_wo[4096];   // exclusive popcount prefix per word (vw <= 4096)
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (t == 0) {
        uint32_t lo = 0, hi = F;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_f = lo;
    }
    __syncthreads();
    const uint32_t f = s_f;
    const uint64_t s0 = (uint64_t)HS[f] + (uint64_t)(wi - wstart[f]) * CH;
    const uint64_t s1 = min((uint64_t)HS[f + 1], s0 + CH);
    const uint32_t nw = (uint32_t)((s1 - s0 + PD_WAVE - 1) / PD_WAVE);
    const uint64_t* vb = valid + (uint64_t)wi * vw;
    // block exclusive scan of the words' popcounts (16 words per thread)
    uint32_t c[16], sum = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const uint32_t j = t * 16 + q;
        c[q] = j < nw ? pd::popc64(vb[j]) : 0u;
        sum += c[q];
    }
    const uint32_t inc = pd::wave_incl_scan(sum);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (uint32_t i = 0; i < w; i++) run += s_w[i];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        s_wo[t * 16 + q] = run;
        run += c[q];
    }
    __syncthreads();
    const uint64_t base = woff[wi];
    const uint64_t lt = pd::lanemask_lt();
    for (uint32_t j = w; j < nw; j += 4) {
        const uint64_t bits = vb[j];
        if ((bits >> l) & 1) {
            const uint64_t src = s0 + (uint64_t)j * PD_WAVE + l;
            const uint64_t dst = base + s_wo[j] + pd::popc64(bits & lt);
            if (dst < out_cap) {
                out_p[dst] = in_p[src];
                out_b[dst] = in_b[src];
            }
        }
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

uint64_t hash_join_global(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                          const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap,
                          hipStream_t s);

static uint32_t next_pow2(uint64_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Largest partition count of the partitioned join (20 hash bits: three
// onesweep passes; partition bounds and work items are 32-bit words).
constexpr uint32_t HJ_MAX_FANOUT = 1u << 20;
// Build rows one LDS table takes (7/8 of its slots).
constexpr uint64_t HJ_LDS_ROWS = (uint64_t)HJ_TS * 7 / 8;

// Partition count of the partitioned join for nb build rows: knob
// hj_fanout, else about hj_rows_per_part build rows per partition; by
// default 2048 rows (a quarter of an LDS table, short probe chains) when
// that needs no more 8-bit partition passes than 4096 rows would, else
// 4096.  At most HJ_MAX_FANOUT.
static uint32_t hj_fanout(pimdal_ctx* c, uint64_t nb) {
    const int64_t tuned = c->tune_get("hj_fanout", 0);
    const int64_t rpp = c->tune_get("hj_rows_per_part", 0);
    auto fan = [&](uint64_t rows) {
        return std::min<uint32_t>(HJ_MAX_FANOUT, std::max<uint32_t>(2, next_pow2(std::max<uint64_t>(1, ceil_div(nb, rows)))));
    };
    auto passes = [](uint32_t f) { return ((uint32_t)__builtin_ctz(f) + 7) / 8; };
    if (tuned > 0) return std::min<uint32_t>(HJ_MAX_FANOUT, std::max<uint32_t>(2, next_pow2((uint64_t)tuned)));
    if (rpp > 0) return fan((uint64_t)rpp);
    return passes(fan(2048)) == passes(fan(4096)) ? fan(2048) : fan(4096);
}

// Returns the number of matching pairs; the first min(count, cap) are in
// out_p / out_b.
//
// A build partition larger than an LDS table (HJ_LDS_ROWS rows: heavy
// duplicate build keys, which every partition count puts in one partition)
// is not refused: the LDS kernels skip it, and after them it is joined on
// its own through one global table in HBM (hash_join_global over the
// partition's build and probe ranges), its pairs appended after the LDS
// partitions' pairs.
uint64_t hash_join(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                   const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap, hipStream_t s) {
    PD_REQUIRE(nb < (1ull << 32) && np < (1ull << 32), PIMDAL_E_INVALID, "join inputs must have < 2^32 rows");
    PD_REQUIRE((bk || !nb) && (pk || !np), PIMDAL_E_INVALID, "key pointer is NULL");
    if (nb == 0 || np == 0) return 0;
    // hj_mode: 1 = one global table (small build sides, whose table the
    // caches hold), 0 = radix-partitioned LDS tables; default by build size
    const int64_t mode = c->tune_get("hj_mode", -1);
    if (mode == 1 || (mode < 0 && nb <= (uint64_t)c->tune_get("hj_global_max_build", 1 << 16)))
        return hash_join_global(c, bk, bp, nb, pk, pp, np, out_p, out_b, cap, s);
    PD_REQUIRE(nb < (1ull << 30) && np < (1ull << 30), PIMDAL_E_INVALID,
               "partitioned hash join supports up to 2^30-1 rows per side");
    uint32_t F = hj_fanout(c, nb);
    // probe rows per work item: 1.5x the average probe partition, so a
    // typical partition is one work item (one LDS table build) while a
    // skewed one is split over several blocks; at most np / (4 x CUs) rows,
    // so there are always enough work items to fill the chip
    const int64_t ch_tuned = c->tune_get("hj_chunk", 0);
    uint32_t CH = (uint32_t)ch_tuned;
    if (ch_tuned <= 0) {
        CH = std::max<uint32_t>(8192, next_pow2(ceil_div(np, (uint64_t)F) * 3 / 2));
        while (CH > 8192 && (uint64_t)CH * 4 * (uint64_t)c->cu_count > np) CH >>= 1;
    }
    // probe (knob hj_probe): 1 = direct output slots + match bitmap (items of
    // at most 262144 rows: the packing kernel scans 4096 bitmap words per
    // item), 0 = compacting rounds
    const bool direct_probe = c->tune_get("hj_probe", 1) == 1;
    if (direct_probe) CH = std::min<uint32_t>(std::max<uint32_t>(CH, 1), 262144);
    const uint32_t vw = (uint32_t)(ceil_div(CH, 8192) * 128);   // bitmap words per item (whole 8192-row rounds)
    constexpr uint32_t GR = 1, GS = 1;   // partition f = [H[f], H[f + 1])
    Scratch flags(c, 16, s);
    const bool direct = cap >= np && out_p && out_b;
    Scratch tmp_p(c, direct ? 0 : np * 4, s), tmp_b(c, direct ? 0 : np * 4, s);
    uint32_t* p0 = direct ? out_p : tmp_p.as<uint32_t>();
    uint32_t* b0 = direct ? out_b : tmp_b.as<uint32_t>();
    // the direct probe's pairs sit at their probe rows' positions, a bitmap
    // per work item marking them: packed unless every probe row matched
    auto finish_direct = [&](const uint32_t* HSp, uint32_t Fx, uint32_t CHx, const uint32_t* wstp, uint32_t* wcntp,
                             uint32_t* woffp, const uint64_t* validp, uint32_t vwx, uint64_t nwi,
                             uint64_t total) -> uint64_t {
        if (direct && total == np) return total;   // dense as written
        exclusive_scan_u32(c, wcntp, woffp, nwi, nullptr, s);
        if (out_p && out_b && cap) {
            if (direct) {
                Scratch cp(c, total * 4 + 4, s), cb(c, total * 4 + 4, s);
                launch(c, "hj_compact", k_hj_compact_bits, dim3((uint32_t)nwi), dim3(256), 0, s, HSp, Fx, CHx, wstp,
                       (const uint32_t*)woffp, validp, vwx, (const uint32_t*)p0, (const uint32_t*)b0,
                       cp.as<uint32_t>(), cb.as<uint32_t>(), total);
                copy_d2d(c, out_p, cp.get(), std::min(total, cap) * 4, s);
                copy_d2d(c, out_b, cb.get(), std::min(total, cap) * 4, s);
                PD_HIP(hipStreamSynchronize(s));
            } else {
                launch(c, "hj_compact", k_hj_compact_bits, dim3((uint32_t)nwi), dim3(256), 0, s, HSp, Fx, CHx, wstp,
                       (const uint32_t*)woffp, validp, vwx, (const uint32_t*)p0, (const uint32_t*)b0, out_p, out_b,
                       cap);
            }
        }
        return total;
    };
    // One partition pass per side (partition.hip) into at most 1024
    // partitions of about 16384 build rows, probed through bucketed LDS
    // tables (k_hj_csr): knob hj_onepass (default 1; off when hj_fanout or
    // hj_rows_per_part is set).  Duplicate build keys or a partition past
    // the table fall through to the general path below.
    const uint32_t F1 = std::max<uint32_t>(2, next_pow2(ceil_div(nb, 16384)));
    if (direct_probe && F1 <= 1024 && c->tune_get("hj_onepass", 1) == 1 && c->tune_get("hj_fanout", 0) <= 0 &&
        c->tune_get("hj_rows_per_part", 0) <= 0) {
        // probe rows per work item: the average partition plus 15 %, in whole
        // 8192-row rounds (a typical partition is one item: one table
        // build), halved while the items would be fewer than 4 per CU
        uint32_t CH1 = (uint32_t)ch_tuned;
        if (ch_tuned <= 0) {
            const uint64_t avg = ceil_div(np, (uint64_t)F1);
            CH1 = (uint32_t)std::min<uint64_t>(262144, std::max<uint64_t>(8192, ceil_div(avg * 23 / 20, 8192) * 8192));
            while (CH1 > 8192 && (uint64_t)F1 * ceil_div(avg, (uint64_t)CH1) < 4 * (uint64_t)c->cu_count)
                CH1 = (uint32_t)std::max<uint64_t>(8192, ceil_div(CH1 / 2, 8192) * 8192);
        }
        CH1 = std::min<uint32_t>(std::max<uint32_t>(CH1, 1), 262144);
        const int64_t hc_rows = c->tune_get("hj_csr_rows", 8);
        PD_REQUIRE(hc_rows == 8 || hc_rows == 16, PIMDAL_E_INVALID, "hj_csr_rows must be 8 or 16");
        const uint64_t round1 = (uint64_t)HC_BLOCK * hc_rows;   // probe rows per block round
        const uint32_t vw1 = (uint32_t)(ceil_div(CH1, round1) * (round1 / 64));
        const uint32_t bits1 = (uint32_t)__builtin_ctz(F1);
        Scratch HR(c, ((uint64_t)F1 + 1) * 4, s), HS(c, ((uint64_t)F1 + 1) * 4, s);
        Scratch R2(c, nb * 8, s), S2(c, np * 8, s);   // partitioned (key, payload) rows
        // the build side's partition pass on the side stream, beside the
        // probe side's (knob hj_overlap, default 1; not while capturing)
        const bool overlap = c->tune_get("hj_overlap", 1) == 1 && !c->capture;
        hipStream_t sb = overlap ? side_fork(c, s) : s;
        SideJoin build_done(c, s, overlap);
        hash_partition_1p(c, bk, bp ? bp : bk, nb, bits1, R2.as<uint2>(), HR.as<uint32_t>(), sb);
        hash_partition_1p(c, pk, pp ? pp : pk, np, bits1, S2.as<uint2>(), HS.as<uint32_t>(), s);
        Scratch wst(c, ((uint64_t)F1 + 1) * 4, s);
        const uint64_t max_wi = F1 + ceil_div(np, CH1);
        Scratch wcnt(c, (max_wi + 1) * 4, s), woff(c, (max_wi + 1) * 4, s);
        launch(c, "hj_chunks", k_hj_plan, dim3(1), dim3(1024), 0, s, (const uint32_t*)HS.as<uint32_t>(), F1, CH1,
               wst.as<uint32_t>(), flags.as<uint32_t>(), wcnt.as<uint32_t>(), (uint32_t)(max_wi + 1));
        Scratch valid(c, max_wi * vw1 * 8, s);
        build_done.join();
        if (hc_rows == 16)
            launch(c, "hj_join", k_hj_csr<16>, dim3((uint32_t)max_wi), dim3(HC_BLOCK), 0, s, (const uint2*)R2.as<uint2>(),
                   (const uint32_t*)HR.as<uint32_t>(), (const uint2*)S2.as<uint2>(), (const uint32_t*)HS.as<uint32_t>(),
                   F1, CH1, (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), p0, b0, valid.as<uint64_t>(),
                   vw1, flags.as<uint32_t>());
        else
            launch(c, "hj_join", k_hj_csr<8>, dim3((uint32_t)max_wi), dim3(HC_BLOCK), 0, s, (const uint2*)R2.as<uint2>(),
                   (const uint32_t*)HR.as<uint32_t>(), (const uint2*)S2.as<uint2>(), (const uint32_t*)HS.as<uint32_t>(),
                   F1, CH1, (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), p0, b0, valid.as<uint64_t>(),
                   vw1, flags.as<uint32_t>());
        uint32_t fl[4];
        read_small(c, fl, flags.get(), 16, s);
        if (!fl[0] && !fl[1])
            return finish_direct(HS.as<uint32_t>(), F1, CH1, wst.as<uint32_t>(), wcnt.as<uint32_t>(),
                                 woff.as<uint32_t>(), valid.as<uint64_t>(), vw1, max_wi,
                                 (uint64_t)fl[2] | ((uint64_t)fl[3] << 32));
        fill_u32(c, flags.get(), 0, 4, s);
    }
    Scratch rk(c, nb * 4, s), rp(c, nb * 4, s), sk(c, np * 4, s), sp(c, np * 4, s);
    for (int attempt = 0;; attempt++) {
        const uint32_t bits = (uint32_t)__builtin_ctz(F);
        Scratch HR(c, ((uint64_t)F + 1) * 4, s), HS(c, ((uint64_t)F + 1) * 4, s);
        hash_partition(c, bk, bp ? bp : bk, 4, nb, bits, rk.as<uint32_t>(), rp.get(), HR.as<uint32_t>(), s);
        hash_partition(c, pk, pp ? pp : pk, 4, np, bits, sk.as<uint32_t>(), sp.get(), HS.as<uint32_t>(), s);
        Scratch wst(c, ((uint64_t)F + 1) * 4, s);
        const uint64_t max_wi = F + ceil_div(np, CH);
        Scratch wcnt(c, (max_wi + 1) * 4, s), woff(c, (max_wi + 1) * 4, s);
        launch(c, "hj_chunks", k_hj_plan, dim3(1), dim3(1024), 0, s, (const uint32_t*)HS.as<uint32_t>(), F, CH,
               wst.as<uint32_t>(), flags.as<uint32_t>(), wcnt.as<uint32_t>(), (uint32_t)(max_wi + 1));
        Scratch valid(c, direct_probe ? max_wi * vw * 8 : 0, s);
        uint32_t fl[4] = {1, 0, 0, 0};   // {duplicate build keys, oversized build partition, pairs (u64)}
        if (direct_probe) {
            launch(c, "hj_join", k_hj_direct<1024, 8>, dim3((uint32_t)max_wi), dim3(1024), 0, s,
                   (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(),
                   (const uint32_t*)HR.as<uint32_t>(), (const uint32_t*)sk.as<uint32_t>(),
                   (const uint32_t*)sp.as<uint32_t>(), (const uint32_t*)HS.as<uint32_t>(), F, CH,
                   (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), p0, b0, valid.as<uint64_t>(), vw,
                   flags.as<uint32_t>());
            read_small(c, fl, flags.get(), 16, s);
            if (!fl[0] && !fl[1])
                return finish_direct(HS.as<uint32_t>(), F, CH, wst.as<uint32_t>(), wcnt.as<uint32_t>(),
                                     woff.as<uint32_t>(), valid.as<uint64_t>(), vw, max_wi,
                                     (uint64_t)fl[2] | ((uint64_t)fl[3] << 32));
            // duplicate build keys or an oversized partition: the counting path below
        } else {
#define PD_HJ_UNIQUE(B, R)                                                                                     \
    launch(c, "hj_join", k_hj_join<HJ_UNIQUE, B, R>, dim3((uint32_t)max_wi), dim3(B), 0, s,                        \
           (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(), (const uint32_t*)HR.as<uint32_t>(), \
           GR, (const uint32_t*)sk.as<uint32_t>(), (const uint32_t*)sp.as<uint32_t>(),                             \
           (const uint32_t*)HS.as<uint32_t>(), GS, F, CH, (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), \
           (const uint32_t*)nullptr, p0, b0, (uint64_t)np, flags.as<uint32_t>())
            if (c->tune_get("hj_block", 1024) == 1024) PD_HJ_UNIQUE(1024, 8);
            else PD_HJ_UNIQUE(512, 16);
#undef PD_HJ_UNIQUE
            read_small(c, fl, flags.get(), 16, s);
            if (!fl[0] && !fl[1]) {
                const uint64_t total = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32);
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
        // ---- counting path: every pair of the partitions an LDS table holds ----
        fill_u32(c, wcnt.get(), 0, max_wi + 1, s);
        fill_u32(c, flags.get(), 0, 4, s);
        launch(c, "hj_join_count", k_hj_join<HJ_COUNT>, dim3((uint32_t)max_wi), dim3(HJ_BLOCK), 0, s,
               (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(),
               (const uint32_t*)HR.as<uint32_t>(), GR, (const uint32_t*)sk.as<uint32_t>(),
               (const uint32_t*)sp.as<uint32_t>(), (const uint32_t*)HS.as<uint32_t>(), GS, F, CH,
               (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(), (const uint32_t*)nullptr,
               (uint32_t*)nullptr, (uint32_t*)nullptr, (uint64_t)0, flags.as<uint32_t>());
        exclusive_scan_u32(c, wcnt.as<uint32_t>(), woff.as<uint32_t>(), max_wi, nullptr, s);
        read_small(c, fl, flags.get(), 16, s);
        // oversized build partitions: their row ranges, from the bounds
        std::vector<uint32_t> big;
        std::vector<uint32_t> hr, hs;
        if (fl[1]) {
            hr.resize((size_t)F + 1);
            hs.resize((size_t)F + 1);
            PD_HIP(hipMemcpyAsync(hr.data(), HR.get(), ((uint64_t)F + 1) * 4, hipMemcpyDeviceToHost, s));
            PD_HIP(hipMemcpyAsync(hs.data(), HS.get(), ((uint64_t)F + 1) * 4, hipMemcpyDeviceToHost, s));
            PD_HIP(hipStreamSynchronize(s));
            for (uint32_t f = 0; f < F; f++)
                if ((uint64_t)hr[f + 1] - hr[f] > HJ_LDS_ROWS) big.push_back(f);
            // many oversized partitions mean too few partitions for the
            // build side (a tuned fan-out), not a few heavy keys: twice the
            // partitions
            if (big.size() * 64 > F && F < HJ_MAX_FANOUT) {
                F *= 2;
                fill_u32(c, flags.get(), 0, 4, s);
                continue;
            }
        }
        const uint64_t lds_total = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32);
        PD_REQUIRE(lds_total < (1ull << 32), PIMDAL_E_UNSUPPORTED, "join output exceeds 2^32-1 pairs");
        if (out_p && out_b && cap && lds_total)
            launch(c, "hj_join_write", k_hj_join<HJ_WRITE>, dim3((uint32_t)max_wi), dim3(HJ_BLOCK), 0, s,
                   (const uint32_t*)rk.as<uint32_t>(), (const uint32_t*)rp.as<uint32_t>(),
                   (const uint32_t*)HR.as<uint32_t>(), GR, (const uint32_t*)sk.as<uint32_t>(),
                   (const uint32_t*)sp.as<uint32_t>(), (const uint32_t*)HS.as<uint32_t>(), GS, F, CH,
                   (const uint32_t*)wst.as<uint32_t>(), wcnt.as<uint32_t>(),
                   (const uint32_t*)woff.as<uint32_t>(), out_p, out_b, cap, flags.as<uint32_t>());
        // the oversized partitions, each through one global table, appended
        uint64_t total = lds_total;
        for (const uint32_t f : big) {
            const uint64_t b0r = hr[f], nbf = (uint64_t)hr[f + 1] - hr[f];
            const uint64_t s0r = hs[f], npf = (uint64_t)hs[f + 1] - hs[f];
            if (!npf) continue;
            const bool room = out_p && out_b && total < cap;
            const uint64_t k = hash_join_global(c, rk.as<uint32_t>() + b0r, rp.as<uint32_t>() + b0r, nbf,
                                                sk.as<uint32_t>() + s0r, sp.as<uint32_t>() + s0r, npf,
                                                room ? out_p + total : nullptr, room ? out_b + total : nullptr,
                                                room ? cap - total : 0, s);
            total += k;
            PD_REQUIRE(total < (1ull << 32), PIMDAL_E_UNSUPPORTED, "join output exceeds 2^32-1 pairs");
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
    PD_NO_CAPTURE(c);
    const uint64_t k = hash_join(c, build_key, build_payload, n_build, probe_key, probe_payload, n_probe,
                                 out_probe_payload, out_build_payload, out_capacity, c->stream);
    PD_HIP(hipStreamSynchronize(c->stream));
    *out_count = k;
    PD_REQUIRE(k <= out_capacity, PIMDAL_E_CAPACITY, "join output larger than out_capacity (count returned)");
    PD_API_END
}

// ===================== non-partitioned mode (global table) =====================
// The build side goes into one open-addressing table in HBM (slot = key |
// payload << 32, occupancy in a separate bitmap so every key and payload
// value is legal); the probe side streams through it once.  At the
// It pays for small build sides only, whose table the Infinity Cache and the
// L2s hold: a 2^24-row build (2^25 slots, 256 MB) sends nearly every probe
// to HBM for a whole cache line, well behind the partitioned mode.  Output
// is compacted per 4096-row probe tile with a decoupled look-back, so it
// comes out in probe order.

namespace pd {

constexpr int HG_BLOCK = 256;
constexpr int HG_ITEMS = 16;
constexpr int HG_TILE = HG_BLOCK * HG_ITEMS;

__device__ __forceinline__ uint64_t hg_slot(uint32_t key, uint64_t mask) { return pd_mix64_dev(key) & mask; }

__global__ void k_hg_build(const uint32_t* __restrict__ bk, const uint32_t* __restrict__ bp, uint64_t nb,
                           uint64_t* __restrict__ tab, uint32_t* __restrict__ occ, uint64_t mask) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = bk[i];
        const uint32_t p = bp ? bp[i] : k;
        uint64_t h = hg_slot(k, mask);
        while (true) {
            const uint32_t bit = 1u << (h & 31);
            if (!(atomicOr(&occ[h >> 5], bit) & bit)) {
                tab[h] = (uint64_t)k | ((uint64_t)p << 32);
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

// MODE 0: emit the first match of each probe row (and flag rows with more
// than one match); MODE 1: count matches per tile; MODE 2: emit every match
// at the tile's offset (offsets from a scan of MODE 1 counts).
template <int MODE>
__global__ __launch_bounds__(HG_BLOCK) void k_hg_probe(const uint32_t* __restrict__ pk, const uint32_t* __restrict__ pp,
                                                       uint64_t np, const uint64_t* __restrict__ tab,
                                                       const uint32_t* __restrict__ occ, uint64_t mask,
                                                       uint32_t* __restrict__ out_p, uint32_t* __restrict__ out_b,
                                                       uint64_t cap, uint64_t* __restrict__ status,
                                                       uint32_t* __restrict__ tile_ctr, uint64_t ntiles,
                                                       uint32_t* __restrict__ tile_cnt,
                                                       const uint32_t* __restrict__ tile_off,
                                                       uint32_t* __restrict__ flags) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[HG_ITEMS][HG_BLOCK / PD_WAVE];
    __shared__ uint32_t s_total;
    __shared__ uint64_t s_excl;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    while (true) {
        if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= ntiles) break;
        const uint64_t base = tile * HG_TILE;
        // probe rows base + i*BLOCK + t (row order = (i, wave, lane))
        uint32_t key[HG_ITEMS], m[HG_ITEMS], first[HG_ITEMS];
        bool dup = false;
#pragma unroll
        for (int i = 0; i < HG_ITEMS; i++) {
            const uint64_t j = base + (uint64_t)i * HG_BLOCK + t;
            m[i] = 0;
            first[i] = 0;
            key[i] = 0;
            if (j < np) {
                const uint32_t k = pk[j];
                key[i] = k;
                uint64_t h = hg_slot(k, mask);
                while (occ[h >> 5] & (1u << (h & 31))) {
                    const uint64_t e = tab[h];
                    if ((uint32_t)e == k) {
                        if (!m[i]) first[i] = (uint32_t)(e >> 32);
                        m[i]++;
                        if (MODE == 0 && m[i] > 1) break;
                    }
                    h = (h + 1) & mask;
                }
                if (MODE == 0) { dup |= m[i] > 1; m[i] = m[i] ? 1 : 0; }
            }
        }
        if (MODE == 0 && dup) atomicOr(&flags[0], 1u);
        if (MODE == 1) {
            uint32_t c = 0;
#pragma unroll
            for (int i = 0; i < HG_ITEMS; i++) c += m[i];
            c = pd::wave_sum(c);
            if (l == 0) s_cnt[0][w] = c;
            __syncthreads();
            if (t == 0) {
                uint32_t tot = 0;
                for (int ww = 0; ww < HG_BLOCK / PD_WAVE; ww++) tot += s_cnt[0][ww];
                tile_cnt[tile] = tot;
            }
            __syncthreads();
            continue;
        }
        // per (item, wave) counts -> exclusive scan in (item, wave, lane) order
#pragma unroll
        for (int i = 0; i < HG_ITEMS; i++) {
            const uint32_t c = pd::wave_sum(m[i]);
            if (l == 0) s_cnt[i][w] = c;
        }
        __syncthreads();
        if (w == 0) {
            const uint32_t v = (&s_cnt[0][0])[l];
            const uint32_t inc = pd::wave_incl_scan(v);
            (&s_cnt[0][0])[l] = inc - v;
            if (l == PD_WAVE - 1) s_total = inc;
        }
        __syncthreads();
        if (MODE == 0) {
            if (w == 0) {
                const uint64_t e = pd::lookback_single<4>(status, tile, s_total);
                if (l == 0) {
                    s_excl = e;
                    if (tile == ntiles - 1) atomicAdd(reinterpret_cast<unsigned long long*>(flags + 2), (unsigned long long)(e + s_total));
                }
            }
        } else if (t == 0) {
            s_excl = tile_off[tile];
        }
        __syncthreads();
        const uint64_t excl = s_excl;
        // inclusive-within-lane prefix of m over lanes below, per item
#pragma unroll
        for (int i = 0; i < HG_ITEMS; i++) {
            const uint32_t below = pd::wave_incl_scan(m[i]) - m[i];
            if (!m[i]) continue;
            uint64_t o = excl + s_cnt[i][w] + below;
            const uint64_t j = base + (uint64_t)i * HG_BLOCK + t;
            const uint32_t pay = pp ? pp[j] : key[i];
            if (MODE == 0) {
                if (o < cap) { out_p[o] = pay; out_b[o] = first[i]; }
            } else {
                uint64_t h = hg_slot(key[i], mask);
                while (occ[h >> 5] & (1u << (h & 31))) {
                    const uint64_t e = tab[h];
                    if ((uint32_t)e == key[i]) {
                        if (o < cap) { out_p[o] = pay; out_b[o] = (uint32_t)(e >> 32); }
                        o++;
                    }
                    h = (h + 1) & mask;
                }
            }
        }
        __syncthreads();
    }
}

uint64_t hash_join_global(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                          const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap,
                          hipStream_t s) {
    uint64_t slots = 1;
    while (slots < 2 * nb) slots <<= 1;
    const uint64_t mask = slots - 1;
    Scratch tab(c, slots * 8, s), occ(c, std::max<uint64_t>(slots / 8, 4), s);
    fill_u32(c, occ.get(), 0, std::max<uint64_t>(slots / 32, 1), s);
    launch(c, "hg_build", k_hg_build, dim3((uint32_t)std::min<uint64_t>(ceil_div(nb, 256), 16384)), dim3(256), 0, s, bk,
           bp, nb, tab.as<uint64_t>(), occ.as<uint32_t>(), mask);
    const uint64_t ntiles = ceil_div(np, HG_TILE);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, (uint64_t)c->cu_count * 8);
    Scratch st(c, 8 * (ntiles + 1), s), flags(c, 16, s);
    fill_u64(c, st.get(), 0, ntiles + 1, s);
    fill_u32(c, flags.get(), 0, 4, s);
    const bool direct = cap >= np && out_p && out_b;
    Scratch tp(c, direct ? 0 : np * 4, s), tb(c, direct ? 0 : np * 4, s);
    uint32_t* p0 = direct ? out_p : tp.as<uint32_t>();
    uint32_t* b0 = direct ? out_b : tb.as<uint32_t>();
    uint32_t* ctr = reinterpret_cast<uint32_t*>(st.as<uint64_t>() + ntiles);
    launch(c, "hg_probe", k_hg_probe<0>, dim3(grid), dim3(HG_BLOCK), 0, s, pk, pp, np,
           (const uint64_t*)tab.as<uint64_t>(), (const uint32_t*)occ.as<uint32_t>(), mask, p0, b0, (uint64_t)np,
           st.as<uint64_t>(), ctr, ntiles, (uint32_t*)nullptr, (const uint32_t*)nullptr, flags.as<uint32_t>());
    uint32_t fl[4];
    read_small(c, fl, flags.get(), 16, s);
    uint64_t total = (uint64_t)fl[2] | ((uint64_t)fl[3] << 32);
    if (!fl[0]) {
        if (!direct && out_p && out_b && cap) {
            copy_d2d(c, out_p, p0, std::min(total, cap) * 4, s);
            copy_d2d(c, out_b, b0, std::min(total, cap) * 4, s);
            PD_HIP(hipStreamSynchronize(s));
        }
        return total;
    }
    // duplicate build keys: count per tile, scan, write every pair
    Scratch tc(c, (ntiles + 1) * 4, s), to(c, (ntiles + 1) * 4, s), tot(c, 8, s);
    fill_u32(c, ctr, 0, 1, s);
    launch(c, "hg_probe_count", k_hg_probe<1>, dim3(grid), dim3(HG_BLOCK), 0, s, pk, pp, np,
           (const uint64_t*)tab.as<uint64_t>(), (const uint32_t*)occ.as<uint32_t>(), mask, (uint32_t*)nullptr,
           (uint32_t*)nullptr, (uint64_t)0, st.as<uint64_t>(), ctr, ntiles, tc.as<uint32_t>(), (const uint32_t*)nullptr,
           flags.as<uint32_t>());
    exclusive_scan_u32(c, tc.as<uint32_t>(), to.as<uint32_t>(), ntiles, tot.as<uint64_t>(), s);
    total = read_u64(c, tot.get(), s);
    PD_REQUIRE(total < (1ull << 32), PIMDAL_E_UNSUPPORTED, "join output exceeds 2^32-1 pairs");
    if (out_p && out_b && cap) {
        fill_u32(c, ctr, 0, 1, s);
        launch(c, "hg_probe_write", k_hg_probe<2>, dim3(grid), dim3(HG_BLOCK), 0, s, pk, pp, np,
               (const uint64_t*)tab.as<uint64_t>(), (const uint32_t*)occ.as<uint32_t>(), mask, out_p, out_b, cap,
               st.as<uint64_t>(), ctr, ntiles, (uint32_t*)nullptr, (const uint32_t*)to.as<uint32_t>(),
               flags.as<uint32_t>());
        PD_HIP(hipStreamSynchronize(s));
    }
    return total;
}

}  // namespace pd
