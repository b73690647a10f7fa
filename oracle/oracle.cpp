/*
 * oracle.cpp — CPU restatement of PIMDAL's DB-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP operators are
 * compared against, and the CPU baseline timed beside them.  Only tests/,
 * the smoke check and the benchmark's CPU-baseline leg load it; the product
 * library (pimdal_amd/) never links, loads or calls it.
 *
 * Each function restates the algorithm of the reference CPU operators in
 * reference/db_operators/ (and, where the configuration needs more than
 * they carry, of the PIM library under pimdal/), citing file:line.  The
 * restatement fixes the reference quirks that would make it an invalid
 * checker: the select tail drop and the unconditional out-of-bounds write
 * (select.cpp:15-16, 69-73), the keys-only unstable sort (sort.cpp:30-40), the
 * set-based build side that drops duplicate build keys (hash_join.cpp:131-135)
 * and the racy sort-merge merge (sort_merge_join.cpp:81-101).  Payload
 * columns are carried through every operator, and sums are exact in
 * int64 / int128.
 *
 * Exported with C linkage for ctypes.  Built by oracle/Makefile with
 * g++ -O3 -fopenmp (reference/db_operators/CMakeLists.txt:16 uses -O3 -fopenmp).
 */
#include <stdint.h>
#include <string.h>
#include <time.h>
#include <algorithm>
#include <vector>
#include <unordered_map>
#include <omp.h>
#include <parallel/algorithm>

#include "../include/pimdal_gen.h"

namespace {

/* Rows per OpenMP work block (reference/db_operators/select.cpp:15). */
constexpr uint64_t kBlock = 8000;

inline uint64_t nblocks(uint64_t n) { return (n + kBlock - 1) / kBlock; }

/* reference/db_operators/hash_join.cpp:20-25: multiplicative hash with the
 * 64-bit golden-ratio constant, byte-swapped, low 32 bits kept. */
inline uint32_t ref_hash(uint32_t input) {
    uint64_t x = (uint64_t)input * 11400714785074694791ULL;
    return (uint32_t)__builtin_bswap64(x);
}

}  // namespace


/* ------------------------------------------------------------------ */
/* Generators (same definitions as the device generators).            */
/* ------------------------------------------------------------------ */

extern "C" void orc_gen_u32(uint64_t seed, uint32_t col, uint64_t n, uint32_t* out) {
    const uint64_t s = pd_stream(seed, col);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = (uint32_t)(pd_rand(s, i) >> 32);
}

/* Uniform in [lo, hi] inclusive, hi - lo < 2^32. */
extern "C" void orc_gen_range_i64(uint64_t seed, uint32_t col, uint64_t n, int64_t lo, int64_t hi, int64_t* out) {
    const uint64_t s = pd_stream(seed, col);
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = lo + (int64_t)pd_uniform_span(pd_rand(s, i), span);
}

extern "C" void orc_gen_range_i32(uint64_t seed, uint32_t col, uint64_t n, int64_t lo, int64_t hi, int32_t* out) {
    const uint64_t s = pd_stream(seed, col);
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = (int32_t)(lo + (int64_t)pd_uniform_span(pd_rand(s, i), span));
}

extern "C" void orc_gen_range_u8(uint64_t seed, uint32_t col, uint64_t n, int64_t lo, int64_t hi, uint8_t* out) {
    const uint64_t s = pd_stream(seed, col);
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = (uint8_t)(lo + (int64_t)pd_uniform_span(pd_rand(s, i), span));
}

/* out[i] = perm_m(offset + i). */
extern "C" void orc_gen_perm(uint64_t seed, uint32_t col, uint64_t m, uint64_t offset, uint64_t n, uint32_t* out) {
    const uint64_t s = pd_stream(seed, col);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = pd_perm(s, m, offset + i);
}

/* out[i] = perm_m(in[i]) (in place allowed). */
extern "C" void orc_perm_apply(uint64_t seed, uint32_t col, uint64_t m, uint64_t n, const uint32_t* in, uint32_t* out) {
    const uint64_t s = pd_stream(seed, col);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = pd_perm(s, m, in[i]);
}

/* cdf[k-1] = sum_{j=1..k} j^-s, summed left to right. */
extern "C" void orc_zipf_cdf(uint64_t m, double s, double* cdf) {
    double acc = 0.0;
    for (uint64_t k = 1; k <= m; k++) {
        acc += 1.0 / __builtin_pow((double)k, s);
        cdf[k - 1] = acc;
    }
}

extern "C" void orc_gen_zipf(uint64_t seed, uint32_t col, const double* cdf, uint64_t m, uint64_t n, uint32_t* out_rank) {
    const uint64_t st = pd_stream(seed, col);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out_rank[i] = pd_zipf_rank(cdf, m, pd_rand(st, i));
}

extern "C" void orc_fmix32_array(uint64_t n, const uint32_t* in, uint32_t* out) {
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) out[i] = pd_fmix32(in[i]);
}

/* Fact-table orderkey: `groups` distinct values perm_{2^31}(j), each repeated
 * n / groups times, rows placed by a row permutation. */
extern "C" void orc_gen_orderkey(uint64_t seed, uint64_t n, uint64_t groups, int32_t* out) {
    const uint64_t sr = pd_stream(seed, PD_COL_ORDERKEY_ROWS);
    const uint64_t sv = pd_stream(seed, PD_COL_ORDERKEY_VALS);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t g = (uint64_t)pd_perm(sr, n, i) % groups;
        out[i] = (int32_t)pd_perm(sv, 1ULL << 31, g);
    }
}

extern "C" uint32_t orc_hash32(uint32_t x) { return ref_hash(x); }

/* ------------------------------------------------------------------ */
/* Selection: reference/db_operators/select.cpp:33-112.               */
/* Block count (sel_cnt_kernel, 38-58), serial exclusive prefix of    */
/* block counts (86-89), order-preserving block write (sel_out_kernel,*/
/* 60-75).  Predicate lo <= key < hi, the half-open form of           */
/* select.cpp:33-36.  Output: payload and row index of each selected  */
/* row, ascending row order (pimdal/select/sel.c:47-111 keeps order). */
/* ------------------------------------------------------------------ */

template <typename K>
static uint64_t select_impl(const K* key, const uint32_t* pay, uint64_t n, K lo, K hi,
                            uint32_t* out_pay, uint32_t* out_rid) {
    const uint64_t nb = nblocks(n);
    std::vector<uint64_t> counts(nb + 1, 0);
    uint64_t total = 0;
    #pragma omp parallel for reduction(+ : total) schedule(static)
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t e = std::min(n, (b + 1) * kBlock);
        uint64_t c = 0;
        for (uint64_t i = b * kBlock; i < e; i++) c += (lo <= key[i] && key[i] < hi);
        counts[b + 1] = c;
        total += c;
    }
    for (uint64_t b = 0; b < nb; b++) counts[b + 1] += counts[b];
    #pragma omp parallel for schedule(static)
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t e = std::min(n, (b + 1) * kBlock);
        uint64_t o = counts[b];
        for (uint64_t i = b * kBlock; i < e; i++) {
            if (lo <= key[i] && key[i] < hi) {
                if (out_pay) out_pay[o] = pay[i];
                if (out_rid) out_rid[o] = (uint32_t)i;
                o++;
            }
        }
    }
    return total;
}

extern "C" uint64_t orc_select_u32(const uint32_t* key, const uint32_t* pay, uint64_t n, uint32_t lo, uint32_t hi,
                        uint32_t* out_pay, uint32_t* out_rid) {
    return select_impl<uint32_t>(key, pay, n, lo, hi, out_pay, out_rid);
}

extern "C" uint64_t orc_select_i32(const int32_t* key, const uint32_t* pay, uint64_t n, int32_t lo, int32_t hi,
                        uint32_t* out_pay, uint32_t* out_rid) {
    return select_impl<int32_t>(key, pay, n, lo, hi, out_pay, out_rid);
}

extern "C" uint64_t orc_select_i64(const int64_t* key, const uint32_t* pay, uint64_t n, int64_t lo, int64_t hi,
                        uint32_t* out_pay, uint32_t* out_rid) {
    return select_impl<int64_t>(key, pay, n, lo, hi, out_pay, out_rid);
}

/* ------------------------------------------------------------------ */
/* Sort: reference/db_operators/sort.cpp:30-40 (copy, then             */
/* __gnu_parallel::sort).  Made stable for (key, payload) pairs by    */
/* sorting (key << 32 | row) words, so ties keep input order, then    */
/* gathering payloads.                                                */
/* ------------------------------------------------------------------ */

template <typename V>
static void sort_pairs_impl(const uint32_t* key, const V* val, uint64_t n, uint32_t* out_key, V* out_val) {
    std::vector<uint64_t> w(n);
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) w[i] = ((uint64_t)key[i] << 32) | i;
    __gnu_parallel::sort(w.begin(), w.end());
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        out_key[i] = (uint32_t)(w[i] >> 32);
        if (out_val) out_val[i] = val[(uint32_t)w[i]];
    }
}

extern "C" void orc_sort_pairs_u32(const uint32_t* key, const uint32_t* val, uint64_t n, uint32_t* out_key, uint32_t* out_val) {
    sort_pairs_impl<uint32_t>(key, val, n, out_key, out_val);
}

extern "C" void orc_sort_pairs_u32_u64(const uint32_t* key, const uint64_t* val, uint64_t n, uint32_t* out_key, uint64_t* out_val) {
    sort_pairs_impl<uint64_t>(key, val, n, out_key, out_val);
}

/* Keys-only sort exactly as reference/db_operators/sort.cpp:30-40. */
extern "C" void orc_sort_keys_u32(const uint32_t* key, uint64_t n, uint32_t* out_key) {
    std::copy(key, key + n, out_key);
    __gnu_parallel::sort(out_key, out_key + n);
}

/* ------------------------------------------------------------------ */
/* Hash join: reference/db_operators/hash_join.cpp:57-219.            */
/* partition() (57-119): per-thread histograms of hash % part_n,      */
/* per-bucket prefix across threads, scatter; order inside a bucket   */
/* is input order.  join_partitions() (121-157): per bucket, build a  */
/* table of the inner rows and probe with the outer rows, writing at  */
/* the prefix of earlier buckets' match counts.  The table here is a  */
/* multimap (open addressing, every build row kept), so duplicate     */
/* build keys produce every matching pair.                            */
/* Output order: bucket, then probe row order, then build row order.  */
/* ------------------------------------------------------------------ */

namespace {

struct Parts {
    std::vector<uint64_t> off;  /* part_n + 1 */
    std::vector<uint32_t> key, pay;
};

Parts partition_rows(const uint32_t* key, const uint32_t* pay, uint64_t n, uint32_t part_n) {
    const int nt = omp_get_max_threads();
    const uint64_t nb = nblocks(n);
    std::vector<uint64_t> hist((size_t)nt * part_n, 0);
    Parts p;
    p.off.assign(part_n + 1, 0);
    p.key.resize(n);
    p.pay.resize(n);
    #pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        uint64_t* h = &hist[(size_t)t * part_n];
        #pragma omp for schedule(static)
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t e = std::min(n, (b + 1) * kBlock);
            for (uint64_t i = b * kBlock; i < e; i++) h[ref_hash(key[i]) % part_n]++;
        }
        #pragma omp single
        {
            uint64_t run = 0;
            for (uint32_t q = 0; q < part_n; q++) {
                p.off[q] = run;
                for (int tt = 0; tt < nt; tt++) {
                    uint64_t c = hist[(size_t)tt * part_n + q];
                    hist[(size_t)tt * part_n + q] = run;
                    run += c;
                }
            }
            p.off[part_n] = run;
        }
        #pragma omp for schedule(static)
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t e = std::min(n, (b + 1) * kBlock);
            for (uint64_t i = b * kBlock; i < e; i++) {
                uint64_t d = h[ref_hash(key[i]) % part_n]++;
                p.key[d] = key[i];
                p.pay[d] = pay ? pay[i] : key[i];
            }
        }
    }
    return p;
}

/* Open-addressing multimap over one bucket's build rows. */
struct BucketTable {
    std::vector<uint32_t> slot;  /* build row + 1, 0 = empty */
    uint32_t mask = 0;
    void build(const uint32_t* k, uint64_t nb) {
        uint64_t cap = 16;
        while (cap < 2 * nb) cap <<= 1;
        slot.assign(cap, 0);
        mask = (uint32_t)(cap - 1);
        for (uint64_t i = 0; i < nb; i++) {
            uint32_t s = (uint32_t)pd_mix64(k[i]) & mask;
            while (slot[s]) s = (s + 1) & mask;
            slot[s] = (uint32_t)i + 1;
        }
    }
};

}  // namespace

/* Returns the number of matching pairs; writes at most `cap` of them. */
extern "C" uint64_t orc_hash_join_u32(const uint32_t* bkey, const uint32_t* bpay, uint64_t nb,
                           const uint32_t* pkey, const uint32_t* ppay, uint64_t np, uint32_t part_n,
                           uint32_t* out_ppay, uint32_t* out_bpay, uint64_t cap) {
    if (part_n == 0) part_n = 16384;  /* hash_join.cpp:163 */
    Parts R = partition_rows(bkey, bpay, nb, part_n);
    Parts S = partition_rows(pkey, ppay, np, part_n);
    std::vector<uint64_t> cnt(part_n + 1, 0);
    /* pass 1: matches per bucket */
    #pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t q = 0; q < part_n; q++) {
        const uint64_t r0 = R.off[q], r1 = R.off[q + 1];
        BucketTable t;
        t.build(&R.key[r0], r1 - r0);
        uint64_t c = 0;
        for (uint64_t j = S.off[q]; j < S.off[q + 1]; j++) {
            uint32_t s = (uint32_t)pd_mix64(S.key[j]) & t.mask;
            while (t.slot[s]) {
                c += (R.key[r0 + t.slot[s] - 1] == S.key[j]);
                s = (s + 1) & t.mask;
            }
        }
        cnt[q + 1] = c;
    }
    for (uint32_t q = 0; q < part_n; q++) cnt[q + 1] += cnt[q];
    /* pass 2: write at the bucket's prefix */
    #pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t q = 0; q < part_n; q++) {
        const uint64_t r0 = R.off[q], r1 = R.off[q + 1];
        BucketTable t;
        t.build(&R.key[r0], r1 - r0);
        uint64_t o = cnt[q];
        for (uint64_t j = S.off[q]; j < S.off[q + 1]; j++) {
            /* collect this probe row's matches in build-row order */
            uint32_t s = (uint32_t)pd_mix64(S.key[j]) & t.mask;
            uint64_t first = o;
            while (t.slot[s]) {
                uint32_t r = t.slot[s] - 1;
                if (R.key[r0 + r] == S.key[j]) {
                    if (o < cap) { out_ppay[o] = S.pay[j]; out_bpay[o] = r; }
                    o++;
                }
                s = (s + 1) & t.mask;
            }
            if (o > first) {
                uint64_t e = std::min(o, cap);
                if (first < e) std::sort(out_bpay + first, out_bpay + e);
                for (uint64_t x = first; x < e; x++) out_bpay[x] = R.pay[r0 + out_bpay[x]];
            }
        }
    }
    return cnt[part_n];
}

/* ------------------------------------------------------------------ */
/* Sort-merge join: reference/db_operators/sort_merge_join.cpp:48-103 */
/* (sort inner, sort outer, merge with a monotone inner cursor) and   */
/* pimdal/join/sort_merge.c:54-83 (the per-block merge).  Corrected:  */
/* inner size read from the inner table, one cursor per block found   */
/* by binary search (no shared cursor), bounds checked before the     */
/* dereference, every duplicate inner key joined, and output written  */
/* densely at the prefix of per-block counts.  Both sides are sorted  */
/* stably, so the output order is (key, outer row, inner row).        */
/* ------------------------------------------------------------------ */

extern "C" uint64_t orc_sort_merge_join_u32(const uint32_t* bkey, const uint32_t* bpay, uint64_t nb,
                                 const uint32_t* pkey, const uint32_t* ppay, uint64_t np,
                                 uint32_t* out_ppay, uint32_t* out_bpay, uint64_t cap) {
    std::vector<uint32_t> rk(nb), rp(nb), sk(np), sp(np);
    sort_pairs_impl<uint32_t>(bkey, bpay, nb, rk.data(), rp.data());
    sort_pairs_impl<uint32_t>(pkey, ppay, np, sk.data(), sp.data());
    const uint64_t nblk = nblocks(np);
    std::vector<uint64_t> cnt(nblk + 1, 0);
    #pragma omp parallel for schedule(static)
    for (uint64_t b = 0; b < nblk; b++) {
        const uint64_t e = std::min(np, (b + 1) * kBlock);
        uint64_t r = std::lower_bound(rk.begin(), rk.end(), sk[b * kBlock]) - rk.begin();
        uint64_t c = 0;
        for (uint64_t j = b * kBlock; j < e; j++) {
            while (r < nb && rk[r] < sk[j]) r++;
            for (uint64_t x = r; x < nb && rk[x] == sk[j]; x++) c++;
        }
        cnt[b + 1] = c;
    }
    for (uint64_t b = 0; b < nblk; b++) cnt[b + 1] += cnt[b];
    #pragma omp parallel for schedule(static)
    for (uint64_t b = 0; b < nblk; b++) {
        const uint64_t e = std::min(np, (b + 1) * kBlock);
        uint64_t r = std::lower_bound(rk.begin(), rk.end(), sk[b * kBlock]) - rk.begin();
        uint64_t o = cnt[b];
        for (uint64_t j = b * kBlock; j < e; j++) {
            while (r < nb && rk[r] < sk[j]) r++;
            for (uint64_t x = r; x < nb && rk[x] == sk[j]; x++) {
                if (o < cap) { out_ppay[o] = sp[j]; out_bpay[o] = rp[x]; }
                o++;
            }
        }
    }
    return cnt[nblk];
}

/* ------------------------------------------------------------------ */
/* Hash aggregation: reference/db_operators/aggregate.cpp:40-151.     */
/* Thread-local maps over 8000-row blocks (40-67), merged under a     */
/* critical section (69-80), gathered into (key, sum) (85-93).  Sums  */
/* are int64 here (the reference sums into int32) and a row count is  */
/* added; the output is in ascending key order (the reference emits   */
/* map iteration order, which is unspecified).                        */
/* ------------------------------------------------------------------ */

extern "C" uint64_t orc_hash_aggr_u32_i64(const uint32_t* key, const int64_t* val, uint64_t n,
                               uint32_t* out_key, int64_t* out_sum, int64_t* out_cnt, uint64_t cap) {
    struct Acc { int64_t sum; int64_t cnt; };
    std::unordered_map<uint32_t, Acc> global;
    const uint64_t nb = nblocks(n);
    #pragma omp parallel
    {
        std::unordered_map<uint32_t, Acc> local;
        #pragma omp for schedule(static)
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t e = std::min(n, (b + 1) * kBlock);
            for (uint64_t i = b * kBlock; i < e; i++) {
                Acc& a = local[key[i]];
                a.sum += val[i];
                a.cnt += 1;
            }
        }
        #pragma omp critical
        {
            for (auto& kv : local) {
                Acc& a = global[kv.first];
                a.sum += kv.second.sum;
                a.cnt += kv.second.cnt;
            }
        }
    }
    std::vector<uint32_t> keys;
    keys.reserve(global.size());
    for (auto& kv : global) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    for (uint64_t g = 0; g < keys.size() && g < cap; g++) {
        out_key[g] = keys[g];
        out_sum[g] = global[keys[g]].sum;
        out_cnt[g] = global[keys[g]].cnt;
    }
    return keys.size();
}

/* ------------------------------------------------------------------ */
/* Fused filter + group-by with the TPC-H Q1 measures:                */
/*   pimdal/main/tpc_h/query1/dpu/kernel_q1_1.c:59-167 (filter, then  */
/*   disc_price = price*(100-disc) at 142, charge = disc_price*(100+  */
/*   tax) at 155, count 1 at 162), kernel_q1_2.c:37-46 (per-group sum */
/*   of every measure) and host_q1.cpp:279-330 (the Acero plan).      */
/* Predicate: date <= cutoff.  Groups: (flag1, flag2), key value      */
/* flag1 * 256 + flag2; or an int32 key column.  sum_charge is int128.*/
/* Output in ascending key order.                                     */
/* ------------------------------------------------------------------ */

struct Q1Acc {
    int64_t qty = 0, price = 0, disc = 0, disc_price = 0;
    __int128 charge = 0;
    int64_t cnt = 0;
};

static uint64_t q1_emit(std::unordered_map<int32_t, Q1Acc>& m, uint64_t cap, int32_t* okey, int64_t* sq,
                        int64_t* sp, int64_t* sd, int64_t* sdp, int64_t* sc /* 2 words per group */, int64_t* cnt) {
    std::vector<int32_t> keys;
    keys.reserve(m.size());
    for (auto& kv : m) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    for (uint64_t g = 0; g < keys.size() && g < cap; g++) {
        const Q1Acc& a = m[keys[g]];
        okey[g] = keys[g];
        sq[g] = a.qty; sp[g] = a.price; sd[g] = a.disc; sdp[g] = a.disc_price;
        sc[2 * g] = (int64_t)(uint64_t)a.charge;
        sc[2 * g + 1] = (int64_t)(a.charge >> 64);
        cnt[g] = a.cnt;
    }
    return keys.size();
}

static uint64_t q1_impl(const int32_t* date, const uint8_t* f1, const uint8_t* f2, const int32_t* okeyc,
                        const int64_t* qty, const int64_t* price, const int64_t* disc, const int64_t* tax,
                        uint64_t n, int32_t cutoff, uint64_t cap, int32_t* okey, int64_t* sq, int64_t* sp,
                        int64_t* sd, int64_t* sdp, int64_t* sc, int64_t* cnt) {
    std::unordered_map<int32_t, Q1Acc> global;
    const uint64_t nb = nblocks(n);
    #pragma omp parallel
    {
        std::unordered_map<int32_t, Q1Acc> local;
        #pragma omp for schedule(static)
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t e = std::min(n, (b + 1) * kBlock);
            for (uint64_t i = b * kBlock; i < e; i++) {
                if (!(date[i] <= cutoff)) continue;
                const int32_t k = okeyc ? okeyc[i] : (int32_t)f1[i] * 256 + (int32_t)f2[i];
                Q1Acc& a = local[k];
                const int64_t dp = price[i] * (100 - disc[i]);
                a.qty += qty[i];
                a.price += price[i];
                a.disc += disc[i];
                a.disc_price += dp;
                a.charge += (__int128)dp * (100 + tax[i]);
                a.cnt += 1;
            }
        }
        #pragma omp critical
        {
            for (auto& kv : local) {
                Q1Acc& a = global[kv.first];
                a.qty += kv.second.qty; a.price += kv.second.price; a.disc += kv.second.disc;
                a.disc_price += kv.second.disc_price; a.charge += kv.second.charge; a.cnt += kv.second.cnt;
            }
        }
    }
    return q1_emit(global, cap, okey, sq, sp, sd, sdp, sc, cnt);
}

extern "C" uint64_t orc_q1_flags(const int32_t* date, const uint8_t* f1, const uint8_t* f2, const int64_t* qty,
                      const int64_t* price, const int64_t* disc, const int64_t* tax, uint64_t n, int32_t cutoff,
                      uint64_t cap, int32_t* okey, int64_t* sq, int64_t* sp, int64_t* sd, int64_t* sdp,
                      int64_t* sc, int64_t* cnt) {
    return q1_impl(date, f1, f2, nullptr, qty, price, disc, tax, n, cutoff, cap, okey, sq, sp, sd, sdp, sc, cnt);
}

extern "C" uint64_t orc_q1_key(const int32_t* date, const int32_t* key, const int64_t* qty, const int64_t* price,
                    const int64_t* disc, const int64_t* tax, uint64_t n, int32_t cutoff, uint64_t cap,
                    int32_t* okey, int64_t* sq, int64_t* sp, int64_t* sd, int64_t* sdp, int64_t* sc, int64_t* cnt) {
    return q1_impl(date, nullptr, nullptr, key, qty, price, disc, tax, n, cutoff, cap, okey, sq, sp, sd, sdp, sc, cnt);
}

/* ------------------------------------------------------------------ */
/* TPC-H Q6 shape:                                                     */
/*   pimdal/main/tpc_h/query6/dpu/kernel_q6.c:91-101 (pred_date:       */
/*   start <= date < end; pred_discount: D-1 <= disc <= D+1;           */
/*   pred_quantity: qty < Q), 113-141 (the three selections in turn,   */
/*   then price * discount and an int64 sum) and host_q6.cpp:46-49     */
/*   (D = 6, Q = 24, one year of dates), 60-66 (sum over devices),     */
/*   69-96 (the Acero plan: the same filters, sum(discount * price)).  */
/* Sums are int64 modulo 2^64, as the reference's int64 accumulators.  */
/* Returns the number of rows kept.                                    */
/* ------------------------------------------------------------------ */
extern "C" uint64_t orc_q6(const int32_t* date, const int64_t* qty, const int64_t* price, const int64_t* disc,
                           uint64_t n, int32_t date_lo, int32_t date_hi, int64_t disc_lo, int64_t disc_hi,
                           int64_t qty_lt, int64_t* revenue) {
    uint64_t rev = 0, cnt = 0;
    const uint64_t nb = nblocks(n);
    #pragma omp parallel for schedule(static) reduction(+ : rev, cnt)
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t e = std::min(n, (b + 1) * kBlock);
        for (uint64_t i = b * kBlock; i < e; i++) {
            if (!(date[i] >= date_lo && date[i] < date_hi)) continue;
            if (!(qty[i] < qty_lt)) continue;
            if (!(disc[i] >= disc_lo && disc[i] <= disc_hi)) continue;
            rev += (uint64_t)disc[i] * (uint64_t)price[i];
            cnt += 1;
        }
    }
    *revenue = (int64_t)rev;
    return cnt;
}

/* ------------------------------------------------------------------ */
/* Sort-based aggregation: pimdal/aggregate/aggregate.c:31-158 (run-  */
/* length reduce of key-sorted input).  Sort by key (stable), then one*/
/* (key, sum, count) per run, ascending key.                          */
/* ------------------------------------------------------------------ */

extern "C" uint64_t orc_sort_aggr_u32_i64(const uint32_t* key, const int64_t* val, uint64_t n,
                               uint32_t* out_key, int64_t* out_sum, int64_t* out_cnt, uint64_t cap) {
    std::vector<uint32_t> sk(n);
    std::vector<uint64_t> sv(n);
    sort_pairs_impl<uint64_t>(key, (const uint64_t*)val, n, sk.data(), sv.data());
    uint64_t g = 0;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i;
        int64_t s = 0;
        while (j < n && sk[j] == sk[i]) { s += (int64_t)sv[j]; j++; }
        if (g < cap) { out_key[g] = sk[i]; out_sum[g] = s; out_cnt[g] = (int64_t)(j - i); }
        g++;
        i = j;
    }
    return g;
}

/* ------------------------------------------------------------------ */
/* Packed dates: pimdal/main/tpc_h/reader/read_table.cpp:8-16         */
/* (date_to_int).  The reference fills a struct tm with strptime of a */
/* "YYYY-MM-DD" literal and packs tm_mday + ((tm_mon + 1) << 5) +     */
/* (tm_year << 9).  The same fields come from the C library here:     */
/* strptime for literals, gmtime_r for date32 values (days since      */
/* 1970-01-01), then the same packing, summed modulo 2^32 (what the   */
/* reference's int arithmetic gives for years too far from 1900).     */
/* ------------------------------------------------------------------ */

static int32_t pack_tm(const struct tm& tm) {
    return (int32_t)((uint32_t)tm.tm_mday + ((uint32_t)(tm.tm_mon + 1) << 5) + ((uint32_t)tm.tm_year << 9));
}

/* Returns 0 and the packed date, or -1 when strptime rejects the text. */
extern "C" int orc_date_pack_str(const char* s, int32_t* out) {
    struct tm tm;
    memset(&tm, 0, sizeof tm);
    const char* end = strptime(s, "%Y-%m-%d", &tm);
    if (!end) return -1;
    *out = pack_tm(tm);
    return 0;
}

extern "C" void orc_date32_pack(const int32_t* days, uint64_t n, int32_t* out) {
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        const time_t t = (time_t)days[i] * 86400;
        struct tm tm;
        gmtime_r(&t, &tm);
        out[i] = pack_tm(tm);
    }
}

/* ------------------------------------------------------------------ */
/* Bloom-filter semi-join of the PIM library: partition by hash0 bits */
/* (pimdal/join/hash_join.c:154-235 part_kernel; bucket =            */
/* (hash0(key) >> shift) & (buckets - 1) per level, MSB first, so the */
/* final partition is bits [shift - part_bits, shift) of hash0),      */
/* one 16384-bit filter per partition set by hash1..hash4 (filter_    */
/* kernel, hash_join.c:117-152; insert_bloom_filter,                  */
/* pimdal/hash/hash_steps.c:143-152), and a probe that keeps an       */
/* element when all four bits are set (match_kernel, hash_join.c:237- */
/* 341; probe_bloom_filter, hash_steps.c:161-183).  The mixers are    */
/* pimdal/hash/hash_func.h:5-58.                                      */
/* The mixers are pinned against hash_func.h itself, compiled into   */
/* oracle/_ref by oracle/ref.mk (tests/test_ref_pins.py).             */
/* ------------------------------------------------------------------ */

static uint32_t pim_mix(uint32_t key, uint32_t add, int s1, int s2, int s3, int s4) {
    key += add;
    key += ~(key << s1);
    key ^= (key >> s2);
    key += (key << s3);
    key ^= (key >> s4);
    return key;
}
static uint32_t pim_hash0(uint32_t k) { return pim_mix(k, 272333, 5, 18, 11, 17); }
static uint32_t pim_bloom_hash(int f, uint32_t k) {
    switch (f) {
        case 0: return pim_mix(k, 593627, 22, 26, 10, 28);   /* hash1 */
        case 1: return pim_mix(k, 274277, 13, 20, 3, 9);     /* hash2 */
        case 2: return pim_mix(k, 855467, 4, 21, 5, 16);     /* hash3 */
        default: return pim_mix(k, 176521, 10, 25, 5, 24);   /* hash4 */
    }
}
constexpr uint32_t kFilterBits = 16384;   /* hash_steps.c:15 FILTER_SIZE */

static uint32_t pim_part(uint32_t key, uint32_t part_bits, uint32_t shift) {
    return part_bits ? (pim_hash0(key) >> (shift - part_bits)) & ((1u << part_bits) - 1) : 0;
}

/* filters: (1 << part_bits) * 2048 bytes, written whole. */
extern "C" void orc_bloom_build(const uint32_t* keys, uint64_t n, uint32_t part_bits, uint32_t shift,
                                uint8_t* filters) {
    const uint64_t fb = kFilterBits / 8;
    memset(filters, 0, ((uint64_t)1 << part_bits) * fb);
    for (uint64_t i = 0; i < n; i++) {
        uint8_t* f = filters + pim_part(keys[i], part_bits, shift) * fb;
        for (int h = 0; h < 4; h++) {
            const uint32_t bit = pim_bloom_hash(h, keys[i]) % kFilterBits;
            f[bit / 8] |= (uint8_t)(1u << (bit % 8));
        }
    }
}

/* Kept probe elements in probe order: payload and row index; returns the count. */
extern "C" uint64_t orc_bloom_probe(const uint32_t* keys, const uint32_t* pay, uint64_t n, uint32_t part_bits,
                                    uint32_t shift, const uint8_t* filters, uint32_t* out_pay, uint32_t* out_row) {
    const uint64_t fb = kFilterBits / 8;
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t* f = filters + pim_part(keys[i], part_bits, shift) * fb;
        bool hit = true;
        for (int h = 0; h < 4 && hit; h++) {
            const uint32_t bit = pim_bloom_hash(h, keys[i]) % kFilterBits;
            hit = (f[bit / 8] >> (bit % 8)) & 1;
        }
        if (hit) {
            if (out_pay) out_pay[k] = pay ? pay[i] : keys[i];
            if (out_row) out_row[k] = (uint32_t)i;
            k++;
        }
    }
    return k;
}

extern "C" uint32_t orc_pim_hash(int which, uint32_t key) { return which == 0 ? pim_hash0(key) : pim_bloom_hash(which - 1, key); }

/* ------------------------------------------------------------------ */
/* TPC-H Q3 shape.                                                     */
/* ------------------------------------------------------------------ */

/* pimdal/main/tpc_h/query3/dpu/kernel_q3_1.c:116-137 (customers of one
 * segment), kernel_q3_2.c:146-190 (orders with o_orderdate < date),
 * kernel_q3_3.c:160-215 (join on custkey), kernel_q3_4.c:120-141 (lineitems
 * with l_shipdate > date), kernel_q3_5.c:66-120, 170-262 (join on orderkey,
 * revenue = extendedprice * (100 - discount) summed per (orderkey,
 * orderdate, shippriority), sorted by revenue) and host_q3.cpp:300-345 (the
 * top rows).  The segment is a dictionary code (the reference compares 16-
 * byte strings); order: revenue descending, then o_orderdate, then orderkey
 * ascending (the reference's quicksort leaves ties in no fixed order).
 * Orderkeys must be unique (o_orderkey is the orders key); returns -1 when
 * one repeats.  Writes at most cap groups; returns the group count. */
extern "C" int64_t orc_q3(const int32_t* c_key, const uint8_t* c_seg, uint64_t nc, const int32_t* o_key,
                          const int32_t* o_cust, const int32_t* o_date, const int32_t* o_prio, uint64_t no,
                          const int32_t* l_key, const int64_t* l_price, const int64_t* l_disc, const int32_t* l_ship,
                          uint64_t nl, uint32_t segment, int32_t date, uint64_t cap, int32_t* out_key,
                          int64_t* out_rev, int32_t* out_date, int32_t* out_prio) {
    std::unordered_map<int32_t, char> cust;
    cust.reserve(nc / 4 + 16);
    for (uint64_t i = 0; i < nc; i++)
        if (c_seg[i] == segment) cust.emplace(c_key[i], 1);
    std::unordered_map<int32_t, uint64_t> order_row;   // qualifying orderkey -> order row
    for (uint64_t i = 0; i < no; i++) {
        if (!(o_date[i] < date) || !cust.count(o_cust[i])) continue;
        if (!order_row.emplace(o_key[i], i).second) return -1;
    }
    std::unordered_map<uint64_t, int64_t> rev;   // order row -> revenue (int64, wrapping like the reference)
    std::unordered_map<uint64_t, char> seen;
    for (uint64_t i = 0; i < nl; i++) {
        if (!(l_ship[i] > date)) continue;
        auto it = order_row.find(l_key[i]);
        if (it == order_row.end()) continue;
        const uint64_t r = it->second;
        rev[r] = (int64_t)((uint64_t)rev[r] + (uint64_t)(l_price[i] * (100 - l_disc[i])));
        seen[r] = 1;
    }
    std::vector<uint64_t> rows;
    rows.reserve(seen.size());
    for (const auto& kv : seen) rows.push_back(kv.first);
    std::sort(rows.begin(), rows.end(), [&](uint64_t a, uint64_t b) {
        if (rev[a] != rev[b]) return rev[a] > rev[b];
        if (o_date[a] != o_date[b]) return o_date[a] < o_date[b];
        return (uint32_t)o_key[a] < (uint32_t)o_key[b];
    });
    for (uint64_t j = 0; j < rows.size() && j < cap; j++) {
        const uint64_t r = rows[j];
        out_key[j] = o_key[r];
        out_rev[j] = rev[r];
        out_date[j] = o_date[r];
        out_prio[j] = o_prio[r];
    }
    return (int64_t)rows.size();
}

/* ------------------------------------------------------------------ */
/* TPC-H Q4 shape.                                                     */
/* ------------------------------------------------------------------ */

/* pimdal/main/tpc_h/query4/dpu/kernel_q4_1.c:161-164 (orders with
 * date_start <= o_orderdate < date_end), kernel_q4_2.c:104-106 (lineitems
 * with l_commitdate < l_receiptdate), kernel_q4_3.c:144-209 (join on
 * orderkey, one row per order), kernel_q4_4.c:72-99 and
 * host_q4.cpp:211-231 (orders counted per o_orderpriority, summed over the
 * devices, ordered by priority).  The priority is a dictionary code (the
 * reference groups 16-byte strings).  Orderkeys of the qualifying orders
 * must be unique; returns -1 when one repeats.  count[code] = qualifying
 * orders of that priority with at least one qualifying lineitem; returns
 * the number of codes with a count. */
extern "C" int64_t orc_q4(const int32_t* o_key, const int32_t* o_date, const uint8_t* o_prio, uint64_t no,
                          const int32_t* l_key, const int32_t* l_commit, const int32_t* l_receipt, uint64_t nl,
                          int32_t date_start, int32_t date_end, int64_t* count /* [256] */) {
    std::unordered_map<int32_t, uint64_t> order_row;   // qualifying orderkey -> order row
    for (uint64_t i = 0; i < no; i++) {
        if (!(o_date[i] >= date_start && o_date[i] < date_end)) continue;
        if (!order_row.emplace(o_key[i], i).second) return -1;
    }
    std::unordered_map<uint64_t, char> exists;
    for (uint64_t i = 0; i < nl; i++) {
        if (!(l_commit[i] < l_receipt[i])) continue;
        auto it = order_row.find(l_key[i]);
        if (it != order_row.end()) exists[it->second] = 1;
    }
    for (int c = 0; c < 256; c++) count[c] = 0;
    for (const auto& kv : exists) count[o_prio[kv.first]]++;
    int64_t g = 0;
    for (int c = 0; c < 256; c++) g += count[c] ? 1 : 0;
    return g;
}

/* ------------------------------------------------------------------ */
/* TPC-H Q5 shape.                                                     */
/* ------------------------------------------------------------------ */

/* pimdal/main/tpc_h/query5/dpu/kernel_q5_1.c:210-212 (the region by name,
 * then its nations and their customers), kernel_q5_2.c:123-146 (orders with
 * date_start <= o_orderdate < date_end), kernel_q5_3.c (customers join
 * orders on custkey), kernel_q5_5.c (join lineitems on orderkey),
 * kernel_q5_7.c:64-100 (join suppliers on suppkey and nationkey, revenue =
 * l_extendedprice * (100 - l_discount)), host_q5.cpp:459-500 (summed per
 * nation, revenue descending; ties here by nationkey).  The region name is
 * a dictionary code.  Returns -1 when a qualifying custkey or orderkey, or a
 * suppkey, repeats; writes at most cap rows; returns the nation count. */
extern "C" int64_t orc_q5(const int32_t* r_key, const uint8_t* r_name, uint64_t nr, const int32_t* n_key,
                          const int32_t* n_reg, uint64_t nn, const int32_t* c_key, const int32_t* c_nat, uint64_t nc,
                          const int32_t* o_key, const int32_t* o_cust, const int32_t* o_date, uint64_t no,
                          const int32_t* l_key, const int32_t* l_supp, const int64_t* l_price, const int64_t* l_disc,
                          uint64_t nl, const int32_t* s_key, const int32_t* s_nat, uint64_t ns, uint32_t region_name,
                          int32_t date_start, int32_t date_end, uint64_t cap, int32_t* out_nat, int64_t* out_rev) {
    std::unordered_map<int32_t, char> nations;
    for (uint64_t i = 0; i < nn; i++)
        for (uint64_t r = 0; r < nr; r++)
            if (r_name[r] == region_name && r_key[r] == n_reg[i]) nations.emplace(n_key[i], 1);
    std::unordered_map<int32_t, int32_t> cust;   // custkey -> nation
    for (uint64_t i = 0; i < nc; i++)
        if (nations.count(c_nat[i]) && !cust.emplace(c_key[i], c_nat[i]).second) return -1;
    std::unordered_map<int32_t, int32_t> ord;    // orderkey -> customer's nation
    for (uint64_t i = 0; i < no; i++) {
        if (!(o_date[i] >= date_start && o_date[i] < date_end)) continue;
        auto it = cust.find(o_cust[i]);
        if (it == cust.end()) continue;
        if (!ord.emplace(o_key[i], it->second).second) return -1;
    }
    std::unordered_map<int32_t, int32_t> supp;   // suppkey -> nation
    for (uint64_t i = 0; i < ns; i++)
        if (!supp.emplace(s_key[i], s_nat[i]).second) return -1;
    std::unordered_map<int32_t, int64_t> rev;
    for (uint64_t i = 0; i < nl; i++) {
        auto o = ord.find(l_key[i]);
        if (o == ord.end()) continue;
        auto sp = supp.find(l_supp[i]);
        if (sp == supp.end() || sp->second != o->second) continue;
        rev[o->second] = (int64_t)((uint64_t)rev[o->second] + (uint64_t)(l_price[i] * (100 - l_disc[i])));
    }
    std::vector<std::pair<int32_t, int64_t>> rows(rev.begin(), rev.end());
    std::sort(rows.begin(), rows.end(), [](const std::pair<int32_t, int64_t>& a, const std::pair<int32_t, int64_t>& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
    });
    for (uint64_t j = 0; j < rows.size() && j < cap; j++) {
        out_nat[j] = rows[j].first;
        out_rev[j] = rows[j].second;
    }
    return (int64_t)rows.size();
}

extern "C" int orc_num_threads(void) { return omp_get_max_threads(); }
extern "C" void orc_set_threads(int t) { omp_set_num_threads(t); }

