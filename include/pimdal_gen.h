/*
 * pimdal_gen.h — counter-based synthetic data generators shared by the host
 * (CPU restatement under oracle/) and the device (pimdal_amd/csrc/gen.hip).
 *
 * Every value is a pure function of (seed, column, row), so the data is the
 * same for any thread count or launch shape.  This replaces the reference's
 * thread-seeded std::default_random_engine generators
 * (reference/db_operators/shared.cpp:9-35) and __gnu_parallel::random_shuffle
 * (reference/db_operators/shared.cpp:37-60), whose output depends on the
 * OpenMP thread count.
 *
 * The functions are header-only and compile both as host C/C++ and as HIP
 * device code (PD_GEN_FN expands to __host__ __device__ under HIP).
 */
#ifndef PIMDAL_GEN_H
#define PIMDAL_GEN_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PD_GEN_FN __host__ __device__ static inline
#else
#define PD_GEN_FN static inline
#endif

/* SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijection on 64 bits. */
PD_GEN_FN uint64_t pd_mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27; x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

/* MurmurHash3 fmix32: a bijection on 32 bits (used to scatter Zipf ranks). */
PD_GEN_FN uint32_t pd_fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85EBCA6BU;
    h ^= h >> 13; h *= 0xC2B2AE35U;
    h ^= h >> 16;
    return h;
}

/* Stream key of one column: distinct columns give unrelated streams. */
PD_GEN_FN uint64_t pd_stream(uint64_t seed, uint32_t col) {
    return pd_mix64(seed + 0x9E3779B97F4A7C15ULL * ((uint64_t)col + 1));
}

/* 64 random bits for row `row` of the stream `stream`. */
PD_GEN_FN uint64_t pd_rand(uint64_t stream, uint64_t row) {
    return pd_mix64(stream ^ pd_mix64(row + 0xD1B54A32D192ED03ULL));
}

/* Uniform integer in [lo, lo + span) for span in [1, 2^32] (multiply-shift). */
PD_GEN_FN uint64_t pd_uniform_span(uint64_t r, uint64_t span) {
    return ((r >> 32) * span) >> 32;
}

/* Uniform double in [0, 1) with 53 random bits. */
PD_GEN_FN double pd_unit_double(uint64_t r) {
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

/* Keyed pseudo-random permutation of [0, m), m in [1, 2^32]: a 4-round
 * balanced Feistel network on the smallest even bit width covering m, with
 * cycle walking back into [0, m).  Bijective for every seed. */
PD_GEN_FN uint32_t pd_perm_bits(uint64_t m) {
    uint32_t b = 2;
    while (b < 32 && (1ULL << b) < m) b++;
    if (b & 1) b++;
    return b;
}

PD_GEN_FN uint64_t pd_feistel(uint64_t stream, uint32_t bits, uint64_t x) {
    const uint32_t h = bits / 2;
    const uint64_t mask = (1ULL << h) - 1;
    uint64_t l = x >> h, r = x & mask;
    for (uint32_t round = 0; round < 4; round++) {
        uint64_t f = pd_mix64(stream + 0x632BE59BD9B4E019ULL * (round + 1) + r) & mask;
        uint64_t nl = r;
        r = l ^ f;
        l = nl;
    }
    return (l << h) | r;
}

PD_GEN_FN uint32_t pd_perm(uint64_t stream, uint64_t m, uint64_t i) {
    const uint32_t bits = pd_perm_bits(m);
    uint64_t x = i;
    do { x = pd_feistel(stream, bits, x); } while (x >= m);
    return (uint32_t)x;
}

/* Zipf(s) ranks over [1, m] by inverse CDF.  `cdf[k-1]` holds
 * sum_{j=1..k} j^-s accumulated left to right in double precision; a rank is
 * the smallest k with cdf[k-1] > u * cdf[m-1]. */
PD_GEN_FN uint32_t pd_zipf_rank(const double* cdf, uint64_t m, uint64_t r) {
    const double target = pd_unit_double(r) * cdf[m - 1];
    uint64_t lo = 0, hi = m - 1;           /* answer index in [lo, hi] */
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (cdf[mid] > target) hi = mid; else lo = mid + 1;
    }
    return (uint32_t)(lo + 1);
}

/* Column ids used by the workload recipes (so host and device agree). */
enum {
    PD_COL_KEY = 1, PD_COL_PAYLOAD = 2, PD_COL_BUILD_PERM = 3, PD_COL_PROBE_FK = 4,
    PD_COL_BUILD_PAY = 5, PD_COL_PROBE_PAY = 6, PD_COL_ZIPF = 7,
    PD_COL_ORDERKEY_ROWS = 10, PD_COL_ORDERKEY_VALS = 11, PD_COL_QTY = 12,
    PD_COL_PRICE = 13, PD_COL_DISC = 14, PD_COL_TAX = 15, PD_COL_DATE = 16,
    PD_COL_FLAG1 = 17, PD_COL_FLAG2 = 18, PD_COL_DIM_PERM = 19, PD_COL_DIM_DATE = 20, PD_COL_DIM_CAT = 21
};

/* Order-independent result checksum terms of the serving layer: a query's
 * checksum is the 64-bit wrapping sum of one term per result row. */
#define PD_CK_SELECT 0x53454C4543540001ULL
#define PD_CK_JOIN 0x4A4F494E00000002ULL
#define PD_CK_GROUP 0x47524F5550000003ULL
/* Selected fact row at output position pos (both < 2^32): the sum over a
 * query's output depends on the order of the rows, not only on the set. */
PD_GEN_FN uint64_t pd_ck_select(uint64_t fact_row, uint64_t pos) {
    return pd_mix64(((pos << 32) | (fact_row & 0xFFFFFFFFull)) ^ PD_CK_SELECT);
}
PD_GEN_FN uint64_t pd_ck_join(uint64_t fact_row, uint64_t dim_row) {
    return pd_mix64(((fact_row << 32) | dim_row) ^ PD_CK_JOIN);
}
PD_GEN_FN uint64_t pd_ck_group(int64_t key, const int64_t* v /* qty, price, disc, disc_price, charge lo, charge hi, count */) {
    uint64_t h = pd_mix64((uint64_t)key ^ PD_CK_GROUP);
    for (int i = 0; i < 7; i++) h = pd_mix64(h ^ (uint64_t)v[i]);
    return h;
}

#endif /* PIMDAL_GEN_H */
