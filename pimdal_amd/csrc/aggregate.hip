// aggregate.hip — hash aggregation, fused filter + group-by with the TPC-H Q1
// measures, and sort-based aggregation.
//
// Reference semantics:
//  * reference/db_operators/aggregate.cpp:40-151 — per-thread hash maps of
//    key -> SUM(val), merged under a critical section, gathered to a table;
//  * pimdal/aggregate/aggregate_hash.c:43-193 — shared WRAM hash table with
//    bucket locks and spill rounds, aggregate applied per element;
//  * pimdal/aggregate/aggregate.c:31-158 — run-length reduce of sorted input;
//  * pimdal/main/tpc_h/query1/dpu/kernel_q1_1.c:138-162 and kernel_q1_2.c:
//    37-46 — Q1 filter, disc_price = price*(100-disc), charge =
//    disc_price*(100+tax), count, per-group sums.
// Every sum is exact: 64-bit integer adds, the charge sum in 128 bits.
//
// Kernels:
//  k_flag_domain    presence bitmaps of the two flag columns (one read of
//                   2 B/row) so group codes can be dense;
//  k_q1_dense<G>    filter + Q1 measures accumulated in registers for G <= 8
//                   dense groups (no atomics), one partial row per block;
//  k_q1_reduce      deterministic 128-bit reduction of the block partials;
//                   the last block to finish writes the result rows;
//  k_q6 / k_q6_final   the TPC-H Q6 filter chain and revenue sum in one
//                   streaming pass (pimdal/main/tpc_h/query6/dpu/kernel_q6.c);
//  k_gb_hist / k_gb_scatter   radix partition of the filtered, projected
//                   rows by the reference hash of the group key;
//  k_gb_lds<Agg>    one partition per block: LDS open-addressing table
//                   (64-bit CAS on (occupied, key) words, so every key value
//                   is legal and no lane ever waits on another), LDS atomic
//                   adds for the measures, then the groups are emitted;
//  k_gb_global<Agg> partitions too large or too diverse for LDS: chunks of
//                   rows insert into a per-partition table in HBM with global
//                   atomics; k_gb_emit_global emits them.
#include <algorithm>
#include <type_traits>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

// ===================== dense (flag) group-by, Q1 measures =====================

constexpr int QD_BLOCK = 256;
constexpr int Q1_NF = 7;  // qty, price, disc, disc_price, charge_lo, charge_hi, count

// Presence bitmaps of the byte values of two flag columns (bits[0..8) for
// flag1, bits[8..16) for flag2).  16 bytes per lane per load; the 256-bit
// masks live in registers, set with unrolled selects (no dynamic register
// indexing).
__device__ __forceinline__ void flag_mark(uint64_t (&m)[4], uint32_t v) {
#pragma unroll
    for (int j = 0; j < 4; j++) m[j] |= ((v >> 6) == (uint32_t)j) ? (1ull << (v & 63)) : 0ull;
}

__global__ __launch_bounds__(256) void k_flag_domain(const uint8_t* __restrict__ f1, const uint8_t* __restrict__ f2,
                                                     uint64_t n, uint32_t* __restrict__ bits /*[16]*/) {
    __shared__ uint32_t s[16];
    if (threadIdx.x < 16) s[threadIdx.x] = 0;
    __syncthreads();
    uint64_t m1[4] = {0, 0, 0, 0}, m2[4] = {0, 0, 0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const bool vec = (((uintptr_t)f1 | (uintptr_t)f2) % 16) == 0;
    const uint64_t n16 = vec ? n / 16 : 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n16; v += stride) {
        const uint4 a = reinterpret_cast<const uint4*>(f1)[v];
        const uint4 b = reinterpret_cast<const uint4*>(f2)[v];
        const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                flag_mark(m1, (wa[k] >> (8 * e)) & 0xFF);
                flag_mark(m2, (wb[k] >> (8 * e)) & 0xFF);
            }
        }
    }
    for (uint64_t i = n16 * 16 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        flag_mark(m1, f1[i]);
        flag_mark(m2, f2[i]);
    }
    // wave OR, then one LDS atomic per word per wave
#pragma unroll
    for (int j = 0; j < 4; j++) {
        for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
            m1[j] |= __shfl_xor(m1[j], d, PD_WAVE);
            m2[j] |= __shfl_xor(m2[j], d, PD_WAVE);
        }
    }
    if (pd::lane() == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (m1[j]) { atomicOr(&s[2 * j], (uint32_t)m1[j]); atomicOr(&s[2 * j + 1], (uint32_t)(m1[j] >> 32)); }
            if (m2[j]) { atomicOr(&s[8 + 2 * j], (uint32_t)m2[j]); atomicOr(&s[8 + 2 * j + 1], (uint32_t)(m2[j] >> 32)); }
        }
    }
    __syncthreads();
    if (threadIdx.x < 16 && s[threadIdx.x]) atomicOr(&bits[threadIdx.x], s[threadIdx.x]);
}

// Builds the code tables from the presence bitmaps: code = rank among the
// present values.  meta[0] = n1, meta[1] = n2.
__global__ void k_flag_lut(const uint32_t* __restrict__ bits, uint8_t* __restrict__ lut /*[512]*/,
                           uint8_t* __restrict__ inv /*[512]*/, uint32_t* __restrict__ meta) {
    if (threadIdx.x != 0) return;
    for (int c = 0; c < 2; c++) {
        uint32_t r = 0;
        for (int v = 0; v < 256; v++) {
            const bool present = (bits[c * 8 + (v >> 5)] >> (v & 31)) & 1;
            lut[c * 256 + v] = present ? (uint8_t)r : (uint8_t)0xFF;   // 0xFF: not in the domain
            if (present) { inv[c * 256 + r] = (uint8_t)v; r++; }
        }
        meta[c] = r;
    }
}

// The code tables of a row sample in one block: presence bitmaps of the
// first `rows` flag bytes, then lut / inv / meta as k_flag_lut builds them
// (code = rank among the present values, in parallel: thread v of column c
// counts the present values below v), and the dense path's miss flag and
// reduction counter cleared (clr[0..1]).
__global__ __launch_bounds__(1024) void k_flag_sample_codes(const uint8_t* __restrict__ f1,
                                                            const uint8_t* __restrict__ f2, uint32_t rows,
                                                            uint8_t* __restrict__ lut /*[512]*/,
                                                            uint8_t* __restrict__ inv /*[512]*/,
                                                            uint32_t* __restrict__ meta, uint32_t* __restrict__ clr) {
    __shared__ uint32_t s[16];
    const uint32_t t = threadIdx.x;
    if (t < 16) s[t] = 0;
    if (t < 2) clr[t] = 0;
    __syncthreads();
    uint64_t m1[4] = {0, 0, 0, 0}, m2[4] = {0, 0, 0, 0};
    const bool vec = (((uintptr_t)f1 | (uintptr_t)f2) % 16) == 0;
    const uint32_t n16 = vec ? rows / 16 : 0;
    for (uint32_t v = t; v < n16; v += 1024) {
        const uint4 a = reinterpret_cast<const uint4*>(f1)[v];
        const uint4 b = reinterpret_cast<const uint4*>(f2)[v];
        const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                flag_mark(m1, (wa[k] >> (8 * e)) & 0xFF);
                flag_mark(m2, (wb[k] >> (8 * e)) & 0xFF);
            }
    }
    for (uint32_t i = n16 * 16 + t; i < rows; i += 1024) {
        flag_mark(m1, f1[i]);
        flag_mark(m2, f2[i]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
            m1[j] |= __shfl_xor(m1[j], d, PD_WAVE);
            m2[j] |= __shfl_xor(m2[j], d, PD_WAVE);
        }
    }
    if (pd::lane() == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (m1[j]) { atomicOr(&s[2 * j], (uint32_t)m1[j]); atomicOr(&s[2 * j + 1], (uint32_t)(m1[j] >> 32)); }
            if (m2[j]) { atomicOr(&s[8 + 2 * j], (uint32_t)m2[j]); atomicOr(&s[8 + 2 * j + 1], (uint32_t)(m2[j] >> 32)); }
        }
    }
    __syncthreads();
    if (t < 512) {
        const uint32_t c = t / 256, v = t % 256;
        uint32_t r = 0;
        for (uint32_t j = 0; j < v / 32; j++) r += __popc(s[c * 8 + j]);
        const uint32_t word = s[c * 8 + v / 32];
        r += __popc(word & ((1u << (v % 32)) - 1u));
        const bool present = (word >> (v % 32)) & 1;
        lut[c * 256 + v] = present ? (uint8_t)r : (uint8_t)0xFF;   // 0xFF: not in the domain
        if (present) inv[c * 256 + r] = (uint8_t)v;
        if (v == 255) meta[c] = r + (present ? 1 : 0);
    }
}

struct Q1Acc {
    int64_t q, p, d, dp, cnt;
    unsigned __int128 ch;
};

template <int G, bool VEC>
__global__ __launch_bounds__(QD_BLOCK) void k_q1_dense(const int32_t* __restrict__ date, const uint8_t* __restrict__ f1,
                                                       const uint8_t* __restrict__ f2, const int64_t* __restrict__ qty,
                                                       const int64_t* __restrict__ price,
                                                       const int64_t* __restrict__ disc,
                                                       const int64_t* __restrict__ tax, uint64_t n, int32_t cutoff,
                                                       const uint8_t* __restrict__ lut,
                                                       const uint32_t* __restrict__ meta,
                                                       uint64_t* __restrict__ partials,
                                                       uint32_t* __restrict__ miss) {
    __shared__ uint8_t s_lut[512];
    __shared__ uint32_t s_miss;
    __shared__ uint64_t s_red[QD_BLOCK / PD_WAVE][G][8];
    for (uint32_t i = threadIdx.x; i < 512; i += QD_BLOCK) s_lut[i] = lut[i];
    if (threadIdx.x == 0) s_miss = 0;
    __syncthreads();
    const uint32_t n2 = meta[1];
    const uint32_t ngr = meta[0] * n2;
    if (ngr > (uint32_t)G) return;  // the general path handles it
    Q1Acc acc[G];
#pragma unroll
    for (int g = 0; g < G; g++) acc[g] = Q1Acc{0, 0, 0, 0, 0, 0};
    bool out_of_domain = false;   // a kept row's flag byte is missing from the code table
    auto add_row = [&](int32_t dt, uint32_t a, uint32_t b, int64_t q, int64_t p, int64_t d, int64_t x) {
        if (dt > cutoff) return;
        const uint32_t ca = s_lut[a], cb = s_lut[256 + b];
        if (ca == 0xFFu || cb == 0xFFu) { out_of_domain = true; return; }
        const uint32_t code = ca * n2 + cb;
        const int64_t dp = p * (100 - d);
        const __int128 ch = (__int128)dp * (100 + x);
#pragma unroll
        for (int g = 0; g < G; g++) {
            if ((uint32_t)g >= ngr) break;   // wave-uniform: only as many slots as codes
            const bool on = code == (uint32_t)g;
            acc[g].q += on ? q : 0;
            acc[g].p += on ? p : 0;
            acc[g].d += on ? d : 0;
            acc[g].dp += on ? dp : 0;
            acc[g].ch += on ? (unsigned __int128)ch : 0;
            acc[g].cnt += on ? 1 : 0;
        }
    };
    const uint64_t stride = (uint64_t)gridDim.x * QD_BLOCK;
    if (VEC) {
        // four consecutive rows per thread per step, 16-byte loads
        const uint64_t nq = n / 4;
        for (uint64_t v = blockIdx.x * (uint64_t)QD_BLOCK + threadIdx.x; v < nq; v += stride) {
            const int4 dt = reinterpret_cast<const int4*>(date)[v];
            const uint32_t fa = reinterpret_cast<const uint32_t*>(f1)[v];
            const uint32_t fb = reinterpret_cast<const uint32_t*>(f2)[v];
            const longlong2 q0 = reinterpret_cast<const longlong2*>(qty)[2 * v], q1 = reinterpret_cast<const longlong2*>(qty)[2 * v + 1];
            const longlong2 p0 = reinterpret_cast<const longlong2*>(price)[2 * v], p1 = reinterpret_cast<const longlong2*>(price)[2 * v + 1];
            const longlong2 d0 = reinterpret_cast<const longlong2*>(disc)[2 * v], d1 = reinterpret_cast<const longlong2*>(disc)[2 * v + 1];
            const longlong2 x0 = reinterpret_cast<const longlong2*>(tax)[2 * v], x1 = reinterpret_cast<const longlong2*>(tax)[2 * v + 1];
            add_row(dt.x, fa & 0xFF, fb & 0xFF, q0.x, p0.x, d0.x, x0.x);
            add_row(dt.y, (fa >> 8) & 0xFF, (fb >> 8) & 0xFF, q0.y, p0.y, d0.y, x0.y);
            add_row(dt.z, (fa >> 16) & 0xFF, (fb >> 16) & 0xFF, q1.x, p1.x, d1.x, x1.x);
            add_row(dt.w, fa >> 24, fb >> 24, q1.y, p1.y, d1.y, x1.y);
        }
        for (uint64_t i = nq * 4 + blockIdx.x * (uint64_t)QD_BLOCK + threadIdx.x; i < n; i += stride)
            add_row(date[i], f1[i], f2[i], qty[i], price[i], disc[i], tax[i]);
    } else {
        for (uint64_t i = blockIdx.x * (uint64_t)QD_BLOCK + threadIdx.x; i < n; i += stride)
            add_row(date[i], f1[i], f2[i], qty[i], price[i], disc[i], tax[i]);
    }
    if (out_of_domain) s_miss = 1;
    // block reduction: wave sums of 64-bit words; the low charge word is
    // split into 32-bit halves so the 64-bit sums cannot drop carries
    const uint32_t w = threadIdx.x / PD_WAVE, l = pd::lane();
#pragma unroll
    for (int g = 0; g < G; g++) {
        uint64_t v[8] = {(uint64_t)acc[g].q, (uint64_t)acc[g].p, (uint64_t)acc[g].d, (uint64_t)acc[g].dp,
                         (uint64_t)acc[g].ch & 0xFFFFFFFFull, (uint64_t)(acc[g].ch >> 64), (uint64_t)acc[g].cnt,
                         ((uint64_t)acc[g].ch) >> 32};
#pragma unroll
        for (int f = 0; f < 8; f++) v[f] = pd::wave_sum64(v[f]);
        if (l == 0)
#pragma unroll
            for (int f = 0; f < 8; f++) s_red[w][g][f] = v[f];
    }
    __syncthreads();
    if (threadIdx.x < G * 8) {
        const uint32_t g = threadIdx.x / 8, f = threadIdx.x % 8;
        uint64_t s = 0;
        for (int ww = 0; ww < QD_BLOCK / PD_WAVE; ww++) s += s_red[ww][g][f];
        partials[((uint64_t)blockIdx.x * G + g) * 8 + f] = s;
    }
    if (threadIdx.x == 0 && s_miss) atomicOr(miss, 1u);
}

// partial layout per (block, group): q, p, d, dp, sum(low 32 bits of the charge's
// low word), sum(charge high word), count, sum(high 32 bits of the charge's low word)
// One block per group code: 128-bit sums of the block partials.
__device__ void q1_final_rows(const unsigned __int128* __restrict__ sums, uint32_t G, const uint32_t* __restrict__ meta,
                              const uint8_t* __restrict__ inv, int32_t* __restrict__ okey, int64_t* __restrict__ sq,
                              int64_t* __restrict__ sp, int64_t* __restrict__ sd, int64_t* __restrict__ sdp,
                              int64_t* __restrict__ sch, int64_t* __restrict__ cnt, uint64_t cap,
                              uint64_t* __restrict__ ngroups);

// One block per group code: 128-bit sums of the block partials.  With
// `done` set, the last block to finish (device-scope counter, zeroed
// beforehand) also writes the result rows (q1_final_rows).
__global__ __launch_bounds__(256) void k_q1_reduce(const uint64_t* __restrict__ partials, uint32_t nblocks, uint32_t G,
                                                   unsigned __int128* __restrict__ sums /*[G][6]*/,
                                                   uint32_t* __restrict__ done, const uint32_t* __restrict__ meta,
                                                   const uint8_t* __restrict__ inv, int32_t* __restrict__ okey,
                                                   int64_t* __restrict__ sq, int64_t* __restrict__ sp,
                                                   int64_t* __restrict__ sd, int64_t* __restrict__ sdp,
                                                   int64_t* __restrict__ sch, int64_t* __restrict__ cnt,
                                                   uint64_t cap, uint64_t* __restrict__ ngroups) {
    __shared__ unsigned __int128 s_v[256][6];
    __shared__ uint32_t s_last;
    const uint32_t g = blockIdx.x, t = threadIdx.x;
    unsigned __int128 a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = t; b < nblocks; b += 256) {
        const uint64_t* r = partials + ((uint64_t)b * G + g) * 8;
#pragma unroll
        for (int f = 0; f < 8; f++) {
            if (f == 4 || f == 6 || f == 7) a[f] += r[f];
            else a[f] += (unsigned __int128)(__int128)(int64_t)r[f];
        }
    }
    s_v[t][0] = a[0]; s_v[t][1] = a[1]; s_v[t][2] = a[2]; s_v[t][3] = a[3];
    s_v[t][4] = a[4] + (a[7] << 32) + (a[5] << 64);
    s_v[t][5] = a[6];
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (t < w)
#pragma unroll
            for (int f = 0; f < 6; f++) s_v[t][f] += s_v[t + w][f];
        __syncthreads();
    }
    if (t < 6) sums[g * 6 + t] = s_v[0][t];
    if (!done) return;
    __threadfence();   // this block's sums before its count
    __syncthreads();
    if (t == 0)
        s_last = __hip_atomic_fetch_add(done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == G - 1;
    __syncthreads();
    if (s_last && t == 0) {
        __threadfence();   // every block's sums before the rows
        q1_final_rows(sums, G, meta, inv, okey, sq, sp, sd, sdp, sch, cnt, cap, ngroups);
    }
}

__device__ void q1_final_rows(const unsigned __int128* __restrict__ sums, uint32_t G, const uint32_t* __restrict__ meta,
                              const uint8_t* __restrict__ inv, int32_t* __restrict__ okey, int64_t* __restrict__ sq,
                              int64_t* __restrict__ sp, int64_t* __restrict__ sd, int64_t* __restrict__ sdp,
                              int64_t* __restrict__ sch, int64_t* __restrict__ cnt, uint64_t cap,
                              uint64_t* __restrict__ ngroups) {
    const uint32_t n1 = meta[0], n2 = meta[1];
    uint64_t k = 0;
    for (uint32_t g = 0; g < G && g < n1 * n2; g++) {
        const unsigned __int128* v = sums + g * 6;
        if (v[5] == 0) continue;
        if (k >= cap) {   // counted, not written: the caller's arrays hold `cap` groups
            k++;
            continue;
        }
        okey[k] = (int32_t)inv[g / n2] * 256 + (int32_t)inv[256 + g % n2];
        sq[k] = (int64_t)v[0];
        sp[k] = (int64_t)v[1];
        sd[k] = (int64_t)v[2];
        sdp[k] = (int64_t)v[3];
        sch[2 * k] = (int64_t)(uint64_t)v[4];
        sch[2 * k + 1] = (int64_t)(uint64_t)(v[4] >> 64);
        cnt[k] = (int64_t)v[5];
        k++;
    }
    *ngroups = k;
}

// ===================== general hash group-by =====================

constexpr int GB_BLOCK = 512;
constexpr uint64_t GB_OCC = 1ull << 32;

// Measure policies.  NI input words per row (after projection), NF output
// fields (u64 each).  fold() turns a projected row into field increments.
struct AggSum {
    static constexpr int NI = 1;   // val
    static constexpr int NF = 2;   // sum, count
    static constexpr int TS = 4096;
    __device__ static void fold(const int64_t* in, uint64_t* f) {
        f[0] = (uint64_t)in[0];
        f[1] = 1;
    }
};

template <int SLOTS>
struct AggQ1T {
    static constexpr int NI = 4;   // qty, price, disc, tax
    static constexpr int NF = 7;   // qty, price, disc, disc_price, charge_lo, charge_hi, count
    static constexpr int TS = SLOTS;
    __device__ static void fold(const int64_t* in, uint64_t* f) {
        const int64_t dp = in[1] * (100 - in[2]);
        const __int128 ch = (__int128)dp * (100 + in[3]);
        f[0] = (uint64_t)in[0];
        f[1] = (uint64_t)in[1];
        f[2] = (uint64_t)in[2];
        f[3] = (uint64_t)dp;
        f[4] = (uint64_t)ch;
        f[5] = (uint64_t)(ch >> 64);
        f[6] = 1;
    }
};

using AggQ1 = AggQ1T<2048>;

// Field add with the 128-bit pair (4, 5) handled by carry for AggQ1.
template <typename Agg>
__device__ __forceinline__ void add_fields(uint64_t* const* col, uint64_t slot, const uint64_t* f) {
    if constexpr (Agg::NF == 7) {
#pragma unroll
        for (int j = 0; j < 7; j++) {
            if (j == 4) {
                const uint64_t old = atomicAdd((unsigned long long*)&col[4][slot], (unsigned long long)f[4]);
                const uint64_t carry = (old + f[4]) < old ? 1 : 0;
                const uint64_t hi = f[5] + carry;
                if (hi) atomicAdd((unsigned long long*)&col[5][slot], (unsigned long long)hi);
            } else if (j != 5) {
                if (f[j]) atomicAdd((unsigned long long*)&col[j][slot], (unsigned long long)f[j]);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < Agg::NF; j++)
            if (f[j]) atomicAdd((unsigned long long*)&col[j][slot], (unsigned long long)f[j]);
    }
}

// Projected row source: key plus NI int64 inputs, with an optional filter.
struct GbSource {
    const uint32_t* key;      // group key (uint32 view); NULL -> flag pair key
    const uint8_t* f1;
    const uint8_t* f2;
    const int32_t* date;      // NULL -> no filter
    int32_t cutoff;
    const int64_t* in[4];
    __device__ bool keep(uint64_t i) const { return !date || date[i] <= cutoff; }
    __device__ uint32_t k(uint64_t i) const { return key ? key[i] : ((uint32_t)f1[i] << 8 | f2[i]); }
};

template <int NI>
struct GbRow {
    uint32_t key, pad;
    int64_t v[NI];
};

// Row sources of the per-partition kernels: load(i) gives the group key and
// the NI int64 inputs of partitioned row i.
template <int NI>
struct RowsAoS {   // GbRow array written by the scatter partitioner
    static constexpr int kNI = NI;
    const GbRow<NI>* rows;
    __device__ bool load(uint64_t i, uint32_t& key, int64_t* in) const {
        const GbRow<NI> r = rows[i];
        key = r.key;
#pragma unroll
        for (int j = 0; j < NI; j++) in[j] = r.v[j];
        return true;
    }
};

struct RowsSum {   // key column + int64 value column (hash aggregation)
    static constexpr int kNI = 1;
    const uint32_t* key;
    const int64_t* val;
    __device__ bool load(uint64_t i, uint32_t& k, int64_t* in) const {
        k = key[i];
        in[0] = val[i];
        return true;
    }
};

// Q1 measures packed into one word: price (int32) in bits 0-31, quantity
// (int16) in 32-47, discount (int8) in 48-55, tax (int8) in 56-63.  The
// quantity -32768 is reserved: it marks a row the date filter dropped.
constexpr uint64_t Q1_SKIP = (uint64_t)0x8000 << 32;
__device__ __forceinline__ uint64_t q1_pack(int64_t qty, int64_t price, int64_t disc, int64_t tax) {
    return (uint64_t)(uint32_t)(int32_t)price | ((uint64_t)(uint16_t)(int16_t)qty << 32) |
           ((uint64_t)(uint8_t)(int8_t)disc << 48) | ((uint64_t)(uint8_t)(int8_t)tax << 56);
}
__device__ __forceinline__ bool q1_packable(int64_t qty, int64_t price, int64_t disc, int64_t tax) {
    return qty == (int16_t)qty && qty != -32768 && price == (int32_t)price && disc == (int8_t)disc &&
           tax == (int8_t)tax;
}

struct RowsQ1Packed {
    static constexpr int kNI = 4;
    const uint32_t* key;
    const uint64_t* w;
    __device__ bool load(uint64_t i, uint32_t& k, int64_t* in) const {
        k = key[i];
        const uint64_t x = w[i];
        in[0] = (int16_t)(x >> 32);
        in[1] = (int32_t)(uint32_t)x;
        in[2] = (int8_t)(x >> 48);
        in[3] = (int8_t)(x >> 56);
        return ((x >> 32) & 0xFFFF) != 0x8000;
    }
};

// Filter + pack for the Q1 group-by on a key column, a pure map (no
// compaction, so no shared counter): row i becomes (key, packed measures),
// or (key, Q1_SKIP) when date > cutoff.  A kept row whose measures do not
// fit the packed widths sets *overflow; the caller then runs the unpacked
// path instead.  Four rows per thread per step (16-byte loads and stores).
template <bool VEC>
__global__ __launch_bounds__(256) void k_q1_pack(const int32_t* __restrict__ date, const uint32_t* __restrict__ key,
                                                 const int64_t* __restrict__ qty, const int64_t* __restrict__ price,
                                                 const int64_t* __restrict__ disc, const int64_t* __restrict__ tax,
                                                 uint64_t n, int32_t cutoff, uint32_t* __restrict__ okey,
                                                 uint64_t* __restrict__ ow, uint32_t* __restrict__ overflow) {
    bool bad = false;
    auto row = [&](int32_t d, int64_t q, int64_t p, int64_t ds, int64_t tx) -> uint64_t {
        if (d > cutoff) return Q1_SKIP;
        bad |= !q1_packable(q, p, ds, tx);
        return q1_pack(q, p, ds, tx);
    };
    const uint64_t n4 = VEC ? n / 4 : 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n4; v += stride) {
        const uint64_t i = v * 4;
        const int4 d = *reinterpret_cast<const int4*>(date + i);
        const uint4 k = *reinterpret_cast<const uint4*>(key + i);
        const longlong2 q0 = *reinterpret_cast<const longlong2*>(qty + i), q1 = *reinterpret_cast<const longlong2*>(qty + i + 2);
        const longlong2 p0 = *reinterpret_cast<const longlong2*>(price + i), p1 = *reinterpret_cast<const longlong2*>(price + i + 2);
        const longlong2 s0 = *reinterpret_cast<const longlong2*>(disc + i), s1 = *reinterpret_cast<const longlong2*>(disc + i + 2);
        const longlong2 t0 = *reinterpret_cast<const longlong2*>(tax + i), t1 = *reinterpret_cast<const longlong2*>(tax + i + 2);
        ulonglong2 w0, w1;
        w0.x = row(d.x, q0.x, p0.x, s0.x, t0.x);
        w0.y = row(d.y, q0.y, p0.y, s0.y, t0.y);
        w1.x = row(d.z, q1.x, p1.x, s1.x, t1.x);
        w1.y = row(d.w, q1.y, p1.y, s1.y, t1.y);
        *reinterpret_cast<uint4*>(okey + i) = k;
        *reinterpret_cast<ulonglong2*>(ow + i) = w0;
        *reinterpret_cast<ulonglong2*>(ow + i + 2) = w1;
    }
    for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        okey[i] = key[i];
        ow[i] = row(date[i], qty[i], price[i], disc[i], tax[i]);
    }
    if (pd::ballot(bad) && pd::lane() == 0) atomicOr(overflow, 1u);
}

// Filter + pack fused with the first partition pass of the Q1 group-by on
// a key column (knob gb_q1_fused, default 1): the rows move to partitions
// of the low lo_bits hash bits as they are packed, so the packed rows are
// written once instead of written by k_q1_pack and read back by the first
// onesweep pass (whose key histogram pass goes too).  8192-row tiles:
//  k_q1p_count    rows per low-bit partition of every tile (the keys
//                 alone: filtered rows are kept, marked Q1_SKIP), into
//                 cnt[tile][1024], and the histogram of the next hash digit
//                 (bits [lo_bits, lo_bits + 8)) for the pass that follows;
//  (partition_tile_scan: per-tile offsets and partition starts)
//  k_q1p_scatter  per tile: the six columns loaded, filtered and packed,
//                 rows ranked in their partition by LDS counters (order
//                 inside a partition does not matter: the next pass is the
//                 stable one), staged in LDS by partition and written in
//                 partition runs; tiles dealt to the XCDs in contiguous runs.
constexpr uint32_t Q1P_BLOCK = 1024;
constexpr uint32_t Q1P_ITEMS = 8;
constexpr uint32_t Q1P_TILE = Q1P_BLOCK * Q1P_ITEMS;   // 8192 rows
constexpr uint32_t Q1P_CSTRIDE = 1024;                 // count row width (partition_tile_scan's)

__global__ __launch_bounds__(512) void k_q1p_count(const uint32_t* __restrict__ key, uint64_t n, uint32_t lo_bits,
                                                   uint32_t* __restrict__ cnt, uint32_t* __restrict__ hist_hi) {
    constexpr uint32_t Q = Q1P_TILE / 512;   // rows per thread
    __shared__ uint32_t h[256];
    __shared__ uint32_t hh[256];
    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint32_t lo_mask = (1u << lo_bits) - 1;
    const uint64_t ntiles = (n + Q1P_TILE - 1) / Q1P_TILE;
    if (t < 256) hh[t] = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if (t < 256) h[t] = 0;
        const uint64_t wave_base = tile * Q1P_TILE + (uint64_t)w * (PD_WAVE * Q);
        const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
            (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * Q), n - wave_base) : 0));
        const __amdgpu_buffer_rsrc_t rk = pd::buf_rsrc(key + wave_base, wrows * 4u);
        uint32_t q[Q];
#pragma unroll
        for (uint32_t i = 0; i < Q; i++) q[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < Q; i++) {
            if (l + i * PD_WAVE < wrows) {
                const uint32_t hv = pd::ref_hash(q[i]);
                atomicAdd(&h[hv & lo_mask], 1u);
                atomicAdd(&hh[(hv >> lo_bits) & 0xFFu], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < Q1P_CSTRIDE; i += 512) cnt[tile * Q1P_CSTRIDE + i] = i <= lo_mask ? h[i] : 0u;
        __syncthreads();
    }
    if (t < 256 && hh[t]) atomicAdd(&hist_hi[t], hh[t]);
}

__global__ __launch_bounds__(Q1P_BLOCK) void k_q1p_scatter(
    const int32_t* __restrict__ date, const uint32_t* __restrict__ key, const int64_t* __restrict__ qty,
    const int64_t* __restrict__ price, const int64_t* __restrict__ disc, const int64_t* __restrict__ tax, uint64_t n,
    int32_t cutoff, uint32_t lo_bits, const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ start, uint32_t xt,
    uint32_t* __restrict__ okey, uint64_t* __restrict__ ow, uint32_t* __restrict__ overflow) {
    __shared__ __align__(16) uint32_t s_k[Q1P_TILE];
    __shared__ __align__(16) uint64_t s_v[Q1P_TILE];
    __shared__ uint32_t s_c[256];   // tile rows per partition, then their start in the tile
    __shared__ uint32_t s_g[256];   // output position of the partition's run minus its tile start
    __shared__ uint32_t s_ws[Q1P_BLOCK / PD_WAVE];
    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint32_t lo_mask = (1u << lo_bits) - 1;
    const uint64_t tile = (uint64_t)(blockIdx.x % 8) * xt + blockIdx.x / 8;
    const uint64_t tile_base = tile * Q1P_TILE;
    if (tile_base >= n) return;   // grid rounded up to a multiple of 8
    const uint32_t tile_rows = (uint32_t)min((uint64_t)Q1P_TILE, n - tile_base);
    uint32_t gpos = 0;
    if (t < 256) {
        s_c[t] = 0;
        gpos = t <= lo_mask ? start[t] + cnt[tile * Q1P_CSTRIDE + t] : 0u;
    }
    // wave w: rows [tile_base + w * 512, + 512), item i = row w * 512 + i * 64 + lane
    const uint64_t wave_base = tile_base + (uint64_t)w * (PD_WAVE * Q1P_ITEMS);
    const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
        (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * Q1P_ITEMS), n - wave_base) : 0));
    const __amdgpu_buffer_rsrc_t rd = pd::buf_rsrc(date + wave_base, wrows * 4u);
    const __amdgpu_buffer_rsrc_t rk = pd::buf_rsrc(key + wave_base, wrows * 4u);
    const __amdgpu_buffer_rsrc_t rq = pd::buf_rsrc(qty + wave_base, wrows * 8u);
    const __amdgpu_buffer_rsrc_t rp = pd::buf_rsrc(price + wave_base, wrows * 8u);
    const __amdgpu_buffer_rsrc_t rs = pd::buf_rsrc(disc + wave_base, wrows * 8u);
    const __amdgpu_buffer_rsrc_t rt = pd::buf_rsrc(tax + wave_base, wrows * 8u);
    // every column's loads issued before any row is packed
    uint32_t kk[Q1P_ITEMS];
    int32_t dd[Q1P_ITEMS];
    int64_t qq[Q1P_ITEMS], pp[Q1P_ITEMS], ss[Q1P_ITEMS], tt[Q1P_ITEMS];
#pragma unroll
    for (uint32_t i = 0; i < Q1P_ITEMS; i++) {
        const uint32_t o = l + i * PD_WAVE;
        kk[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, o * 4u, 0, 0);
        dd[i] = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rd, o * 4u, 0, 0);
        qq[i] = (int64_t)__builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rq, o * 8u, 0, 0));
        pp[i] = (int64_t)__builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rp, o * 8u, 0, 0));
        ss[i] = (int64_t)__builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rs, o * 8u, 0, 0));
        tt[i] = (int64_t)__builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rt, o * 8u, 0, 0));
    }
    uint64_t pw[Q1P_ITEMS];
    bool bad = false;
#pragma unroll
    for (uint32_t i = 0; i < Q1P_ITEMS; i++) {
        if (dd[i] > cutoff) {
            pw[i] = Q1_SKIP;
        } else {
            bad |= l + i * PD_WAVE < wrows && !q1_packable(qq[i], pp[i], ss[i], tt[i]);
            pw[i] = q1_pack(qq[i], pp[i], ss[i], tt[i]);
        }
    }
    __syncthreads();
    uint32_t rk_d[Q1P_ITEMS];   // rank in the partition | partition << 16
#pragma unroll
    for (uint32_t i = 0; i < Q1P_ITEMS; i++) {
        rk_d[i] = 0xFFFFFFFFu;
        if (l + i * PD_WAVE < wrows) {
            const uint32_t d = pd::ref_hash(kk[i]) & lo_mask;
            rk_d[i] = atomicAdd(&s_c[d], 1u) | (d << 16);
        }
    }
    __syncthreads();
    // tile-local partition starts (exclusive scan over the 256 counters, waves 0-3)
    const uint32_t c = t < 256 ? s_c[t] : 0u;
    const uint32_t inc = pd::wave_incl_scan(c);
    if (l == PD_WAVE - 1) s_ws[w] = inc;
    __syncthreads();
    if (t < 256) {
        uint32_t add = 0;
        for (uint32_t i = 0; i < w; i++) add += s_ws[i];
        const uint32_t st = add + inc - c;
        s_c[t] = st;
        s_g[t] = gpos - st;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < Q1P_ITEMS; i++) {
        if (rk_d[i] == 0xFFFFFFFFu) continue;
        const uint32_t j = s_c[rk_d[i] >> 16] + (rk_d[i] & 0xFFFFu);
        s_k[j] = kk[i];
        s_v[j] = pw[i];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < Q1P_ITEMS; i++) {
        const uint32_t j = i * Q1P_BLOCK + t;
        if (j >= tile_rows) break;
        const uint32_t k = s_k[j];
        const uint64_t dst = (uint64_t)s_g[pd::ref_hash(k) & lo_mask] + j;
        okey[dst] = k;
        ow[dst] = s_v[j];
    }
    if (pd::ballot(bad) && l == 0) atomicOr(overflow, 1u);
}

// Filter + pack + the Q1 group-by's hash partition of n rows into 2^bits
// partitions (9 <= bits <= 16): the fused first pass over the low bits - 8
// bits, then hash_partition_finish over the top 8.  (qk, qv): the
// partitioned rows; bounds as hash_partition's; *ovf set when a kept row's
// measures do not fit the packed widths.
static void q1_pack_partition(pimdal_ctx* c, const int32_t* date, const uint32_t* key, const int64_t* qty,
                              const int64_t* price, const int64_t* disc, const int64_t* tax, uint64_t n,
                              int32_t cutoff, uint32_t bits, uint32_t* qk, uint64_t* qv, uint32_t* bounds,
                              uint32_t* ovf, hipStream_t s) {
    PD_REQUIRE(bits >= 9 && bits <= 16 && n >= 1 && n < (1ull << 30), PIMDAL_E_INTERNAL, "fused Q1 partition shape");
    const uint32_t lo_bits = bits - 8;
    const uint32_t ntiles = (uint32_t)ceil_div(n, (uint64_t)Q1P_TILE);
    const uint32_t xt = (ntiles + 7) / 8;
    // cnt[ntiles][1024] | start[1025] | hist of the top digit[256]
    Scratch sc(c, ((uint64_t)ntiles * Q1P_CSTRIDE + 1025 + 256) * 4, s);
    uint32_t* cnt = sc.as<uint32_t>();
    uint32_t* start = cnt + (uint64_t)ntiles * Q1P_CSTRIDE;
    uint32_t* hist_hi = start + 1025;
    fill_u32(c, hist_hi, 0, 256, s);
    const uint32_t cgrid = std::min<uint32_t>(ntiles, (uint32_t)c->cu_count * 4);
    launch(c, "q1_count", k_q1p_count, dim3(cgrid), dim3(512), 0, s, key, n, lo_bits, cnt, hist_hi);
    partition_tile_scan(c, cnt, ntiles, 1u << lo_bits, start, s);
    Scratch pk(c, n * 4, s), pv(c, n * 8, s);
    // (buffer loads: any 4- and 8-byte aligned columns)
    launch(c, "q1_pack", k_q1p_scatter, dim3(8 * xt), dim3(Q1P_BLOCK), 0, s, date, key, qty, price, disc, tax, n,
           cutoff, lo_bits, (const uint32_t*)cnt, (const uint32_t*)start, xt, pk.as<uint32_t>(), pv.as<uint64_t>(),
           ovf);
    hash_partition_finish(c, pk.as<uint32_t>(), pv.as<uint64_t>(), n, bits, lo_bits, hist_hi, qk, qv, bounds, s);
}

// Output policies of the per-partition kernels: put(o, key, fields).
struct OutFields {   // key array + NF u64 field columns of stride `stride`
    uint32_t* key;
    uint64_t* f;
    uint64_t stride;
    template <int NF>
    __device__ void put(uint64_t o, uint32_t k, const uint64_t (&v)[NF]) const {
        key[o] = k;
#pragma unroll
        for (int j = 0; j < NF; j++) f[(uint64_t)j * stride + o] = v[j];
    }
};

struct OutSum {      // hash aggregation: key, sum, count arrays
    uint32_t* key;
    int64_t* sum;
    int64_t* cnt;
    __device__ void put(uint64_t o, uint32_t k, const uint64_t (&v)[2]) const {
        key[o] = k;
        sum[o] = (int64_t)v[0];
        cnt[o] = (int64_t)v[1];
    }
};

struct OutQ1 {       // the Q1 result arrays (pimdal_q1_result)
    int32_t* key;
    int64_t *sq, *sp, *sd, *sdp, *sch, *cnt;
    __device__ void put(uint64_t o, uint32_t k, const uint64_t (&v)[7]) const {
        key[o] = (int32_t)k;
        sq[o] = (int64_t)v[0];
        sp[o] = (int64_t)v[1];
        sd[o] = (int64_t)v[2];
        sdp[o] = (int64_t)v[3];
        sch[2 * o] = (int64_t)v[4];
        sch[2 * o + 1] = (int64_t)v[5];
        cnt[o] = (int64_t)v[6];
    }
};

__global__ __launch_bounds__(512) void k_gb_hist(GbSource src, uint64_t n, uint64_t chunk, uint32_t F,
                                                 uint32_t* __restrict__ H) {
    extern __shared__ uint32_t h[];
    const uint32_t t = threadIdx.x, b = blockIdx.x, G = gridDim.x;
    for (uint32_t f = t; f < F; f += 512) h[f] = 0;
    __syncthreads();
    const uint64_t start = (uint64_t)b * chunk, end = min(n, start + chunk);
    for (uint64_t i = start + t; i < end; i += 512)
        if (src.keep(i)) atomicAdd(&h[pd::ref_hash(src.k(i)) & (F - 1)], 1u);
    __syncthreads();
    for (uint32_t f = t; f < F; f += 512) H[(uint64_t)f * G + b] = h[f];
}

template <int NI>
__global__ __launch_bounds__(512) void k_gb_scatter(GbSource src, uint64_t n, uint64_t chunk, uint32_t F,
                                                    const uint32_t* __restrict__ Hs, GbRow<NI>* __restrict__ out,
                                                    uint64_t cap) {
    extern __shared__ uint32_t ctr[];
    const uint32_t t = threadIdx.x, b = blockIdx.x, G = gridDim.x;
    for (uint32_t f = t; f < F; f += 512) ctr[f] = Hs[(uint64_t)f * G + b];
    __syncthreads();
    const uint64_t start = (uint64_t)b * chunk, end = min(n, start + chunk);
    for (uint64_t i = start + t; i < end; i += 512) {
        if (!src.keep(i)) continue;
        GbRow<NI> r;
        r.key = src.k(i);
        r.pad = 0;
#pragma unroll
        for (int j = 0; j < NI; j++) r.v[j] = src.in[j][i];
        const uint32_t pos = atomicAdd(&ctr[pd::ref_hash(r.key) & (F - 1)], 1u);
        if (pos < cap) out[pos] = r;
    }
}

// gb meta: [0] n_out (u64), [1] any partition sent to the global path.
// One 1024-thread block per partition: the LDS table takes most of a CU's
// LDS, so the block is as wide as allowed to keep 16 waves per CU.
constexpr int GB_LDS_BLOCK = 1024;
template <typename Agg, typename Src, typename Out>
__global__ __launch_bounds__(GB_LDS_BLOCK) void k_gb_lds(Src src, const uint32_t* __restrict__ H, uint32_t G,
                                                         uint32_t F, uint32_t max_rows,
                                                         uint32_t* __restrict__ to_global, Out out,
                                                         uint64_t* __restrict__ meta) {
    constexpr int TS = Agg::TS;
    __shared__ uint64_t s_key[TS];
    __shared__ uint64_t s_f[Agg::NF][TS];
    __shared__ uint32_t s_used, s_over;
    __shared__ uint64_t s_base;
    const uint32_t p = blockIdx.x, t = threadIdx.x;
    const uint64_t r0 = H[(uint64_t)p * G], r1 = H[(uint64_t)(p + 1) * G];
    if (r0 == r1) return;
    if (r1 - r0 > max_rows) {
        if (t == 0) { to_global[p] = 1; atomicOr((uint32_t*)&meta[1], 1u); }
        return;
    }
    for (uint32_t i = t; i < TS; i += GB_LDS_BLOCK) {
        s_key[i] = 0;
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) s_f[j][i] = 0;
    }
    if (t == 0) { s_used = 0; s_over = 0; }
    __syncthreads();
    uint64_t* cols[Agg::NF];
#pragma unroll
    for (int j = 0; j < Agg::NF; j++) cols[j] = s_f[j];
    const uint32_t limit = TS * 3 / 4;
    constexpr int R = 2;   // rows per thread per round, loaded together (a typical partition: one round)
    // Each round inserts its rows' keys, then adds their measures.  After the
    // last round's inserts the group count is final, so the block's output
    // range is claimed (one device atomic) there and its round trip overlaps
    // the last additions.
    unsigned long long base = 0;
    for (uint64_t rb = r0; rb < r1; rb += (uint64_t)GB_LDS_BLOCK * R) {
        uint32_t key[R];
        int64_t in[R][Agg::NI];
        bool ok[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            const uint64_t i = rb + (uint64_t)q * GB_LDS_BLOCK + t;
            ok[q] = i < r1 && src.load(i, key[q], in[q]);
        }
        uint32_t slot[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            slot[q] = TS;   // no slot
            if (!ok[q]) continue;
            const uint64_t want = GB_OCC | key[q];
            uint32_t h = pd::slot_hash(key[q]) & (TS - 1);
            uint32_t probes = 0;
            while (true) {
                const uint64_t cur = s_key[h];
                if (cur == want) break;
                if (cur == 0) {
                    const uint64_t prev = atomicCAS((unsigned long long*)&s_key[h], 0ull, (unsigned long long)want);
                    if (prev == 0) {
                        if (atomicAdd(&s_used, 1u) >= limit) s_over = 1;
                        break;
                    }
                    if (prev == want) break;
                }
                h = (h + 1) & (TS - 1);
                if (++probes >= TS) { s_over = 1; h = TS; break; }
            }
            slot[q] = h;
        }
        const bool last = rb + (uint64_t)GB_LDS_BLOCK * R >= r1;
        if (last) {
            __syncthreads();   // every insert done: s_used and s_over are final
            if (s_over) break;
            if (t == 0) {
                // opaque offset: keeps the compiler from reading the result at once
                uint32_t z = 0;
                asm volatile("" : "+v"(z));
                base = atomicAdd((unsigned long long*)&meta[0] + z, (unsigned long long)s_used);
            }
        }
#pragma unroll
        for (int q = 0; q < R; q++) {
            if (slot[q] >= TS) continue;
            uint64_t f[Agg::NF];
            Agg::fold(in[q], f);
            add_fields<Agg>(cols, slot[q], f);
        }
        // (no early exit on overflow: the barrier above needs every thread;
        // inserts into a full table stop at their probe limit)
    }
    __syncthreads();
    if (s_over) {
        if (t == 0) { to_global[p] = 1; atomicOr((uint32_t*)&meta[1], 1u); }
        return;
    }
    if (t == 0) s_base = base;
    __syncthreads();
    // emit the occupied slots (compaction through an LDS counter)
    if (t == 0) s_used = 0;
    __syncthreads();
    for (uint32_t i = t; i < TS; i += GB_LDS_BLOCK) {
        const uint64_t kw = s_key[i];
        if (!kw) continue;
        const uint64_t o = s_base + atomicAdd(&s_used, 1u);
        uint64_t v[Agg::NF];
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) v[j] = s_f[j][i];
        out.put(o, (uint32_t)kw, v);
    }
}

// work items over partitions routed to HBM tables
__global__ void k_gb_chunks(const uint32_t* __restrict__ H, uint32_t G, uint32_t F, uint32_t CH,
                            const uint32_t* __restrict__ to_global, uint32_t* __restrict__ nch) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < F; p += gridDim.x * blockDim.x) {
        const uint32_t len = H[(uint64_t)(p + 1) * G] - H[(uint64_t)p * G];
        nch[p] = to_global[p] ? (len + CH - 1) / CH : 0;
    }
}

// Table for partition p occupies slots [2*start_p, 2*end_p) of the region:
// capacity twice its row count, so it never fills.
template <typename Agg, typename Src>
__global__ __launch_bounds__(256) void k_gb_global(Src src, const uint32_t* __restrict__ H, uint32_t G, uint32_t F,
                                                   uint32_t CH, const uint32_t* __restrict__ wstart,
                                                   uint64_t* __restrict__ gkey, uint64_t* __restrict__ gf,
                                                   uint64_t gstride) {
    __shared__ uint32_t s_p;
    const uint32_t wi = blockIdx.x;
    if (wi >= wstart[F]) return;
    if (threadIdx.x == 0) {
        uint32_t lo = 0, hi = F;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wstart[mid] <= wi) lo = mid; else hi = mid;
        }
        s_p = lo;
    }
    __syncthreads();
    const uint32_t p = s_p;
    const uint64_t r0 = H[(uint64_t)p * G], r1 = H[(uint64_t)(p + 1) * G];
    const uint64_t c0 = r0 + (uint64_t)(wi - wstart[p]) * CH, c1 = min(r1, c0 + CH);
    const uint64_t base = 2 * r0, cap = 2 * (r1 - r0);
    uint64_t* cols[Agg::NF];
#pragma unroll
    for (int j = 0; j < Agg::NF; j++) cols[j] = gf + (uint64_t)j * gstride + base;
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += 256) {
        uint32_t key;
        int64_t in[Agg::NI];
        if (!src.load(i, key, in)) continue;
        uint64_t f[Agg::NF];
        Agg::fold(in, f);
        const uint64_t want = GB_OCC | key;
        uint64_t h = pd::slot_hash(key) % cap;
        while (true) {
            const uint64_t cur = pd::ld_relaxed(&gkey[base + h]);
            if (cur == want) break;
            if (cur == 0) {
                const uint64_t prev = atomicCAS((unsigned long long*)&gkey[base + h], 0ull, (unsigned long long)want);
                if (prev == 0 || prev == want) break;
            }
            h = (h + 1 == cap) ? 0 : h + 1;
        }
        add_fields<Agg>(cols, h, f);
    }
}

template <typename Agg, typename Out>
__global__ void k_gb_emit_global(const uint32_t* __restrict__ H, uint32_t G, uint32_t F,
                                 const uint32_t* __restrict__ to_global, const uint64_t* __restrict__ gkey,
                                 const uint64_t* __restrict__ gf, uint64_t gstride, Out out,
                                 uint64_t* __restrict__ meta) {
    const uint64_t n2 = 2 * (uint64_t)H[(uint64_t)F * G];
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < n2; s += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t kw = gkey[s];
        if (!kw) continue;
        const uint64_t o = atomicAdd((unsigned long long*)&meta[0], 1ull);
        uint64_t v[Agg::NF];
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) v[j] = gf[(uint64_t)j * gstride + s];
        out.put(o, (uint32_t)kw, v);
    }
}

}  // namespace pd

namespace pd {

// ---- low-cardinality path: per-block LDS tables over large row ranges,
// merged into one small HBM table.  meta[2] = aborted (a block's table or
// the merged table filled up). ----
constexpr uint32_t GB_DIRECT_SLOTS = 65536;

template <typename Agg>
__global__ __launch_bounds__(GB_BLOCK) void k_gb_direct(GbSource src, uint64_t n, uint64_t chunk,
                                                        uint64_t* __restrict__ gkey, uint64_t* __restrict__ gf,
                                                        uint64_t* __restrict__ meta) {
    constexpr int TS = Agg::TS;
    __shared__ uint64_t s_key[TS];
    __shared__ uint64_t s_f[Agg::NF][TS];
    __shared__ uint32_t s_used, s_over;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < TS; i += GB_BLOCK) {
        s_key[i] = 0;
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) s_f[j][i] = 0;
    }
    if (t == 0) { s_used = 0; s_over = 0; }
    __syncthreads();
    uint64_t* cols[Agg::NF];
#pragma unroll
    for (int j = 0; j < Agg::NF; j++) cols[j] = s_f[j];
    const uint32_t limit = TS / 2;
    const uint64_t start = (uint64_t)blockIdx.x * chunk, end = min(n, start + chunk);
    uint32_t iter = 0;
    for (uint64_t i0 = start; i0 < end; i0 += GB_BLOCK) {
        // every 64 rounds, stop early if another block gave up
        if ((iter++ & 63) == 0) {
            if (t == 0 && pd::ld_relaxed(&meta[2])) s_over = 1;
            __syncthreads();
            if (s_over) return;
        }
        const uint64_t i = i0 + t;
        if (i < end && src.keep(i)) {
            int64_t in[Agg::NI];
#pragma unroll
            for (int j = 0; j < Agg::NI; j++) in[j] = src.in[j][i];
            uint64_t f[Agg::NF];
            Agg::fold(in, f);
            const uint32_t key = src.k(i);
            const uint64_t want = GB_OCC | key;
            uint32_t h = pd::slot_hash(key) & (TS - 1);
            bool ok = true;
            uint32_t probes = 0;
            while (true) {
                const uint64_t cur = s_key[h];
                if (cur == want) break;
                if (cur == 0) {
                    const uint64_t prev = atomicCAS((unsigned long long*)&s_key[h], 0ull, (unsigned long long)want);
                    if (prev == 0) {
                        if (atomicAdd(&s_used, 1u) >= limit) s_over = 1;
                        break;
                    }
                    if (prev == want) break;
                }
                h = (h + 1) & (TS - 1);
                if (++probes >= TS ||
                    __hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    s_over = 1;
                    ok = false;
                    break;
                }
            }
            if (ok) add_fields<Agg>(cols, h, f);
        }
        __syncthreads();
        if (s_over) {
            if (t == 0) pd::st_relaxed(&meta[2], (uint64_t)1);
            return;
        }
    }
    __syncthreads();
    // merge into the HBM table
    uint64_t* gcols[Agg::NF];
#pragma unroll
    for (int j = 0; j < Agg::NF; j++) gcols[j] = gf + (uint64_t)j * GB_DIRECT_SLOTS;
    for (uint32_t i = t; i < TS; i += GB_BLOCK) {
        const uint64_t kw = s_key[i];
        if (!kw) continue;
        uint32_t h = pd::slot_hash((uint32_t)kw) & (GB_DIRECT_SLOTS - 1);
        uint32_t probes = 0;
        bool ok = true;
        while (true) {
            const uint64_t cur = pd::ld_relaxed(&gkey[h]);
            if (cur == kw) break;
            if (cur == 0) {
                const uint64_t prev = atomicCAS((unsigned long long*)&gkey[h], 0ull, (unsigned long long)kw);
                if (prev == 0) {
                    if (atomicAdd((unsigned long long*)&meta[3], 1ull) >= GB_DIRECT_SLOTS / 2)
                        pd::st_relaxed(&meta[2], (uint64_t)1);
                    break;
                }
                if (prev == kw) break;
            }
            h = (h + 1) & (GB_DIRECT_SLOTS - 1);
            if (++probes >= GB_DIRECT_SLOTS) { pd::st_relaxed(&meta[2], (uint64_t)1); ok = false; break; }
        }
        if (!ok) continue;
        uint64_t f[Agg::NF];
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) f[j] = s_f[j][i];
        add_fields<Agg>(gcols, h, f);
    }
}

template <typename Agg>
__global__ void k_gb_direct_emit(const uint64_t* __restrict__ gkey, const uint64_t* __restrict__ gf,
                                 uint32_t* __restrict__ out_key, uint64_t* __restrict__ out_f, uint64_t out_stride,
                                 uint64_t* __restrict__ meta) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < GB_DIRECT_SLOTS; s += gridDim.x * blockDim.x) {
        const uint64_t kw = gkey[s];
        if (!kw) continue;
        const uint64_t o = atomicAdd((unsigned long long*)&meta[0], 1ull);
        out_key[o] = (uint32_t)kw;
#pragma unroll
        for (int j = 0; j < Agg::NF; j++) out_f[(uint64_t)j * out_stride + o] = gf[(uint64_t)j * GB_DIRECT_SLOTS + s];
    }
}

static uint32_t gb_next_pow2(uint64_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// k_q1_unpack (below) reshapes staged engine fields into the Q1 arrays.
__global__ void k_q1_unpack(uint64_t g, const uint32_t* __restrict__ key, const uint64_t* __restrict__ f,
                            uint64_t stride, int32_t* __restrict__ okey, int64_t* __restrict__ sq,
                            int64_t* __restrict__ sp, int64_t* __restrict__ sd, int64_t* __restrict__ sdp,
                            int64_t* __restrict__ sch, int64_t* __restrict__ cnt);

// Copies `groups` staged rows (OutFields layout, stride `stride`) to the
// caller's output.
template <int NF>
void finish_staged(pimdal_ctx* c, const OutFields& dst, const uint32_t* sk, const uint64_t* sf, uint64_t stride,
                   uint64_t groups, hipStream_t s) {
    copy_d2d(c, dst.key, sk, groups * 4, s);
    for (int j = 0; j < NF; j++) copy_d2d(c, dst.f + (uint64_t)j * dst.stride, sf + (uint64_t)j * stride, groups * 8, s);
}
template <int NF>
void finish_staged(pimdal_ctx* c, const OutSum& dst, const uint32_t* sk, const uint64_t* sf, uint64_t stride,
                   uint64_t groups, hipStream_t s) {
    copy_d2d(c, dst.key, sk, groups * 4, s);
    copy_d2d(c, dst.sum, sf, groups * 8, s);
    copy_d2d(c, dst.cnt, sf + stride, groups * 8, s);
}
template <int NF>
void finish_staged(pimdal_ctx* c, const OutQ1& dst, const uint32_t* sk, const uint64_t* sf, uint64_t stride,
                   uint64_t groups, hipStream_t s) {
    if (groups)
        launch(c, "q1_unpack", k_q1_unpack, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(groups, 256), 4096))),
               dim3(256), 0, s, groups, sk, sf, stride, dst.key, dst.sq, dst.sp, dst.sd, dst.sdp, dst.sch, dst.cnt);
}

// Per-partition aggregation of partitioned rows (partition p spans
// [H[p*G], H[(p+1)*G])): one LDS table per partition, partitions that do
// not fit go to HBM tables.  Groups are written through `out` when it can
// hold one group per row (cap >= nrows), else staged and copied once the
// count is known.  Returns the group count.
template <typename Agg, typename Src, typename Out>
uint64_t aggregate_parts(pimdal_ctx* c, Src src, const uint32_t* H, uint32_t G, uint32_t F, uint64_t nrows,
                         const Out& out, uint64_t cap, hipStream_t s) {
    Scratch meta(c, 8 * 4, s);
    fill_u64(c, meta.get(), 0, 4, s);
    Scratch tog(c, (uint64_t)F * 4, s);
    fill_u32(c, tog.get(), 0, F, s);
    const uint32_t max_rows = (uint32_t)c->tune_get("gb_lds_max_rows", 65536);
    const bool direct_out = cap >= nrows;
    Scratch sk(c, direct_out ? 0 : nrows * 4, s), sf(c, direct_out ? 0 : nrows * 8 * Agg::NF, s);
    auto run = [&](const auto& o) {
        using O = std::decay_t<decltype(o)>;
        launch(c, "gb_lds", k_gb_lds<Agg, Src, O>, dim3(F), dim3(GB_LDS_BLOCK), 0, s, src, H, G, F, max_rows,
               tog.as<uint32_t>(), o, meta.as<uint64_t>());
        uint64_t m[4];
        read_small(c, m, meta.get(), 32, s);
        if (m[1]) {
            const uint32_t CH = 16384;
            Scratch wst(c, ((uint64_t)F + 1) * 4, s);
            launch(c, "gb_chunks", k_gb_chunks, dim3(std::min<uint32_t>(ceil_div(F, 256), 1024)), dim3(256), 0, s, H, G,
                   F, CH, (const uint32_t*)tog.as<uint32_t>(), wst.as<uint32_t>());
            exclusive_scan_u32(c, wst.as<uint32_t>(), wst.as<uint32_t>(), F, nullptr, s);
            const uint64_t slots = 2 * nrows;
            Scratch gk(c, slots * 8, s), gfv(c, slots * 8 * Agg::NF, s);
            fill_u64(c, gk.get(), 0, slots, s);
            fill_u64(c, gfv.get(), 0, slots * Agg::NF, s);
            const uint64_t max_wi = F + ceil_div(nrows, CH);
            launch(c, "gb_global", k_gb_global<Agg, Src>, dim3((uint32_t)max_wi), dim3(256), 0, s, src, H, G, F, CH,
                   (const uint32_t*)wst.as<uint32_t>(), gk.as<uint64_t>(), gfv.as<uint64_t>(), slots);
            launch(c, "gb_emit_global", k_gb_emit_global<Agg, O>,
                   dim3((uint32_t)std::min<uint64_t>(ceil_div(slots, 256), 4096)), dim3(256), 0, s, H, G, F,
                   (const uint32_t*)tog.as<uint32_t>(), (const uint64_t*)gk.as<uint64_t>(),
                   (const uint64_t*)gfv.as<uint64_t>(), slots, o, meta.as<uint64_t>());
            read_small(c, m, meta.get(), 32, s);
        }
        return m[0];
    };
    if (direct_out) {
        const uint64_t groups = run(out);
        PD_REQUIRE(groups <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
        return groups;
    }
    const OutFields stage{sk.as<uint32_t>(), sf.as<uint64_t>(), nrows};
    const uint64_t groups = run(stage);
    PD_REQUIRE(groups <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    finish_staged<Agg::NF>(c, out, sk.as<uint32_t>(), sf.as<uint64_t>(), nrows, groups, s);
    PD_HIP(hipStreamSynchronize(s));
    return groups;
}

// Partition count for n rows: about rows_target rows per partition.
static uint32_t gb_part_bits(pimdal_ctx* c, uint64_t n, int64_t dflt_rows) {
    const int64_t rows_target = c->tune_get("gb_rows_per_part", dflt_rows);
    uint32_t bits = 1;
    while (bits < 16 && ((uint64_t)1 << bits) * (uint64_t)rows_target < n) bits++;
    return bits;
}

// Group-by over rows already projected to (key column, Src): partitioned
// by hash digits with the onesweep passes (long sequential runs on the
// write side), then aggregate_parts.
template <typename Agg, typename Src, typename MakeSrc, typename Out>
uint64_t group_by_onesweep(pimdal_ctx* c, const uint32_t* key, const void* payload, uint32_t vw, uint64_t n,
                           MakeSrc make_src, const Out& out, uint64_t cap, hipStream_t s) {
    if (n == 0) return 0;
    const uint32_t bits = gb_part_bits(c, n, (int64_t)Agg::TS * 2);
    const uint32_t F = 1u << bits;
    Scratch pk(c, n * 4, s), pv(c, n * vw, s), bounds(c, ((uint64_t)F + 1) * 4, s);
    hash_partition(c, key, payload, vw, n, bits, pk.as<uint32_t>(), pv.get(), bounds.as<uint32_t>(), s);
    return aggregate_parts<Agg>(c, make_src(pk.as<uint32_t>(), pv.get()), bounds.as<uint32_t>(), 1u, F, n, out, cap, s);
}

// Q1 by key, packed rows: filter + pack, partition by hash digits, one LDS
// table per partition.  Fused (knob gb_q1_fused, default 1, partitions of
// 9 to 16 bits): q1_pack_partition; otherwise k_q1_pack into (key, 8-byte
// measures) and hash_partition.  *ovf set when a kept row's measures do not
// fit the packed widths (the outputs are then not valid).
static void q1_pack_rows(pimdal_ctx* c, const int32_t* date, const uint32_t* key, const int64_t* qty,
                         const int64_t* price, const int64_t* disc, const int64_t* tax, uint64_t n, int32_t cutoff,
                         uint32_t bits, uint32_t* qk, uint64_t* qv, uint32_t* bounds, uint32_t* ovf, hipStream_t s) {
    if (bits >= 9 && bits <= 16 && c->tune_get("gb_q1_fused", 1)) {
        q1_pack_partition(c, date, key, qty, price, disc, tax, n, cutoff, bits, qk, qv, bounds, ovf, s);
        return;
    }
    Scratch pk(c, n * 4, s), pw(c, n * 8, s);
    const bool vec = ((uintptr_t)date | (uintptr_t)key | (uintptr_t)qty | (uintptr_t)price | (uintptr_t)disc |
                      (uintptr_t)tax) % 16 == 0;
    const uint32_t gp = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 4), (uint64_t)c->cu_count * 8));
    if (vec)
        launch(c, "q1_pack", k_q1_pack<true>, dim3(gp), dim3(256), 0, s, date, key, qty, price, disc, tax, n, cutoff,
               pk.as<uint32_t>(), pw.as<uint64_t>(), ovf);
    else
        launch(c, "q1_pack", k_q1_pack<false>, dim3(gp), dim3(256), 0, s, date, key, qty, price, disc, tax, n, cutoff,
               pk.as<uint32_t>(), pw.as<uint64_t>(), ovf);
    hash_partition(c, pk.as<uint32_t>(), pw.get(), 8, n, bits, qk, qv, bounds, s);
}

template <typename Agg>
static uint64_t q1_key_packed(pimdal_ctx* c, const int32_t* date, const uint32_t* key, const int64_t* qty,
                              const int64_t* price, const int64_t* disc, const int64_t* tax, uint64_t n, int32_t cutoff,
                              const OutQ1& o, uint64_t cap, uint32_t* ovf, hipStream_t s) {
    if (n == 0) return 0;
    const uint32_t bits = gb_part_bits(c, n, (int64_t)Agg::TS * 2);
    const uint32_t F = 1u << bits;
    Scratch qk(c, n * 4, s), qv(c, n * 8, s), bounds(c, ((uint64_t)F + 1) * 4, s);
    q1_pack_rows(c, date, key, qty, price, disc, tax, n, cutoff, bits, qk.as<uint32_t>(), qv.as<uint64_t>(),
                 bounds.as<uint32_t>(), ovf, s);
    return aggregate_parts<Agg>(c, RowsQ1Packed{qk.as<uint32_t>(), (const uint64_t*)qv.get()}, bounds.as<uint32_t>(),
                                1u, F, n, o, cap, s);
}

// Low-cardinality attempt: per-block LDS tables merged into one small HBM
// table.  Returns false (nothing written) when a table filled up; skipped
// when the group-count hint says the groups cannot fit.
template <typename Agg>
bool group_by_direct(pimdal_ctx* c, const GbSource& src, uint64_t n, uint32_t* out_key, uint64_t* out_f, uint64_t cap,
                     uint64_t* groups, hipStream_t s) {
    const int64_t hint = c->tune_get("gb_groups_hint", 0);
    if (!(hint == 0 || hint <= (int64_t)(Agg::TS / 4))) return false;
    Scratch meta(c, 8 * 4, s);
    fill_u64(c, meta.get(), 0, 4, s);
    Scratch gk(c, 8ull * GB_DIRECT_SLOTS, s), gfv(c, 8ull * GB_DIRECT_SLOTS * Agg::NF, s);
    fill_u64(c, gk.get(), 0, GB_DIRECT_SLOTS, s);
    fill_u64(c, gfv.get(), 0, (uint64_t)GB_DIRECT_SLOTS * Agg::NF, s);
    const uint32_t G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 4096), (uint64_t)c->cu_count * 2));
    const uint64_t chunk = ceil_div(n, G);
    launch(c, "gb_direct", k_gb_direct<Agg>, dim3(G), dim3(GB_BLOCK), 0, s, src, n, chunk, gk.as<uint64_t>(),
           gfv.as<uint64_t>(), meta.as<uint64_t>());
    uint64_t m[4];
    read_small(c, m, meta.get(), 32, s);
    if (m[2]) return false;
    PD_REQUIRE(m[3] <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    launch(c, "gb_direct_emit", k_gb_direct_emit<Agg>, dim3(64), dim3(256), 0, s, (const uint64_t*)gk.as<uint64_t>(),
           (const uint64_t*)gfv.as<uint64_t>(), out_key, out_f, cap, meta.as<uint64_t>());
    PD_HIP(hipStreamSynchronize(s));
    *groups = m[3];
    return true;
}

// Group-by driver over a projected row source (general path: filter in the
// source, rows scattered into partitions as GbRow).  Output: keys in
// out_key, field j of group g at out_f[j * cap + g]; returns the number of
// groups (unspecified order).
template <typename Agg>
uint64_t group_by(pimdal_ctx* c, const GbSource& src, uint64_t n, uint32_t* out_key, uint64_t* out_f, uint64_t cap,
                  hipStream_t s) {
    PD_REQUIRE(n < (1ull << 31), PIMDAL_E_INVALID, "group-by supports up to 2^31-1 rows per call");
    if (n == 0) return 0;
    uint64_t groups = 0;
    if (group_by_direct<Agg>(c, src, n, out_key, out_f, cap, &groups, s)) return groups;
    // partitioned path
    const int64_t rows_target = c->tune_get("gb_rows_per_part", (int64_t)Agg::TS * 2);
    const uint32_t F = std::min<uint32_t>(16384, gb_next_pow2(std::max<uint64_t>(1, ceil_div(n, (uint64_t)rows_target))));
    const uint32_t G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 16384), (uint64_t)c->cu_count * 2));
    const uint64_t chunk = ceil_div(n, G);
    Scratch H(c, ((uint64_t)F * G + 1) * 4, s);
    launch(c, "gb_hist", k_gb_hist, dim3(G), dim3(512), F * 4, s, src, n, chunk, F, H.as<uint32_t>());
    Scratch tot(c, 8, s);
    exclusive_scan_u32(c, H.as<uint32_t>(), H.as<uint32_t>(), (uint64_t)F * G, tot.as<uint64_t>(), s);
    const uint64_t nsel = read_u64(c, tot.get(), s);
    Scratch rows(c, std::max<uint64_t>(nsel, 1) * sizeof(GbRow<Agg::NI>), s);
    launch(c, "gb_scatter", k_gb_scatter<Agg::NI>, dim3(G), dim3(512), F * 4, s, src, n, chunk, F,
           (const uint32_t*)H.as<uint32_t>(), rows.as<GbRow<Agg::NI>>(), nsel);
    return aggregate_parts<Agg>(c, RowsAoS<Agg::NI>{rows.as<GbRow<Agg::NI>>()}, H.as<uint32_t>(), G, F, nsel,
                                OutFields{out_key, out_f, cap}, cap, s);
}

// ---- reshape engine output (u64 fields) into the Q1 result arrays ----
__global__ void k_q1_unpack(uint64_t g, const uint32_t* __restrict__ key, const uint64_t* __restrict__ f,
                            uint64_t stride, int32_t* __restrict__ okey, int64_t* __restrict__ sq,
                            int64_t* __restrict__ sp, int64_t* __restrict__ sd, int64_t* __restrict__ sdp,
                            int64_t* __restrict__ sch, int64_t* __restrict__ cnt) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g; i += (uint64_t)gridDim.x * blockDim.x) {
        okey[i] = (int32_t)key[i];
        sq[i] = (int64_t)f[i];
        sp[i] = (int64_t)f[stride + i];
        sd[i] = (int64_t)f[2 * stride + i];
        sdp[i] = (int64_t)f[3 * stride + i];
        sch[2 * i] = (int64_t)f[4 * stride + i];
        sch[2 * i + 1] = (int64_t)f[5 * stride + i];
        cnt[i] = (int64_t)f[6 * stride + i];
    }
}

// TPC-H Q6 shape (pimdal/main/tpc_h/query6/dpu/kernel_q6.c:91-141): rows
// with date_lo <= date < date_hi, quantity < qty_lt and disc_lo <= discount
// <= disc_hi add price * discount to the revenue.  The reference selects in
// three passes with row-id gathers; here one streaming pass reads the four
// columns (28 bytes per row, four rows per thread with 16-byte loads) and
// keeps the sums in registers.  Sums are modulo 2^64 like the reference's
// int64 accumulators.  Per-block partials: [block][0] revenue, [1] rows.
struct Q6Args {
    int32_t date_lo, date_hi;
    int64_t disc_lo, disc_hi, qty_lt;
};
constexpr int Q6_BLOCK = 256;
template <bool VEC>
__global__ __launch_bounds__(Q6_BLOCK) void k_q6(const int32_t* __restrict__ date, const int64_t* __restrict__ qty,
                                                 const int64_t* __restrict__ price, const int64_t* __restrict__ disc,
                                                 uint64_t n, Q6Args a, uint64_t* __restrict__ partials) {
    __shared__ uint64_t s_red[Q6_BLOCK / PD_WAVE][2];
    uint64_t rev = 0, cnt = 0;
    auto row = [&](int32_t dt, int64_t q, int64_t p, int64_t d) {
        const bool keep = dt >= a.date_lo && dt < a.date_hi && q < a.qty_lt && d >= a.disc_lo && d <= a.disc_hi;
        rev += keep ? (uint64_t)p * (uint64_t)d : 0;
        cnt += keep ? 1 : 0;
    };
    const uint64_t stride = (uint64_t)gridDim.x * Q6_BLOCK;
    if (VEC) {
        const uint64_t nq = n / 4;
        for (uint64_t v = blockIdx.x * (uint64_t)Q6_BLOCK + threadIdx.x; v < nq; v += stride) {
            const int4 dt = reinterpret_cast<const int4*>(date)[v];
            const longlong2 q0 = reinterpret_cast<const longlong2*>(qty)[2 * v], q1 = reinterpret_cast<const longlong2*>(qty)[2 * v + 1];
            const longlong2 p0 = reinterpret_cast<const longlong2*>(price)[2 * v], p1 = reinterpret_cast<const longlong2*>(price)[2 * v + 1];
            const longlong2 d0 = reinterpret_cast<const longlong2*>(disc)[2 * v], d1 = reinterpret_cast<const longlong2*>(disc)[2 * v + 1];
            row(dt.x, q0.x, p0.x, d0.x);
            row(dt.y, q0.y, p0.y, d0.y);
            row(dt.z, q1.x, p1.x, d1.x);
            row(dt.w, q1.y, p1.y, d1.y);
        }
        for (uint64_t i = nq * 4 + blockIdx.x * (uint64_t)Q6_BLOCK + threadIdx.x; i < n; i += stride)
            row(date[i], qty[i], price[i], disc[i]);
    } else {
        for (uint64_t i = blockIdx.x * (uint64_t)Q6_BLOCK + threadIdx.x; i < n; i += stride)
            row(date[i], qty[i], price[i], disc[i]);
    }
    rev = pd::wave_sum64(rev);
    cnt = pd::wave_sum64(cnt);
    const uint32_t w = threadIdx.x / PD_WAVE;
    if (pd::lane() == 0) {
        s_red[w][0] = rev;
        s_red[w][1] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint64_t v = 0;
        for (int ww = 0; ww < Q6_BLOCK / PD_WAVE; ww++) v += s_red[ww][threadIdx.x];
        partials[2 * (uint64_t)blockIdx.x + threadIdx.x] = v;
    }
}

// Sums the block partials (one block; modulo 2^64): out[0] revenue, out[1] rows.
__global__ __launch_bounds__(256) void k_q6_final(const uint64_t* __restrict__ partials, uint32_t nblocks,
                                                  uint64_t* __restrict__ out) {
    __shared__ uint64_t s_red[256 / PD_WAVE][2];
    uint64_t rev = 0, cnt = 0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
        rev += partials[2 * (uint64_t)b];
        cnt += partials[2 * (uint64_t)b + 1];
    }
    rev = pd::wave_sum64(rev);
    cnt = pd::wave_sum64(cnt);
    if (pd::lane() == 0) {
        s_red[threadIdx.x / PD_WAVE][0] = rev;
        s_red[threadIdx.x / PD_WAVE][1] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint64_t v = 0;
        for (int ww = 0; ww < 256 / PD_WAVE; ww++) v += s_red[ww][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Revenue and kept-row count of the Q6 filter chain; synchronises the stream.
uint64_t q6(pimdal_ctx* c, const int32_t* date, const int64_t* qty, const int64_t* price, const int64_t* disc,
            uint64_t n, const Q6Args& a, int64_t* revenue, hipStream_t s) {
    PD_REQUIRE((date && qty && price && disc) || n == 0, PIMDAL_E_INVALID, "missing input column");
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, Q6_BLOCK * 4 * 4),
                                                                           (uint64_t)c->cu_count * 8));
    Scratch part(c, (uint64_t)nb * 16, s), res(c, 16, s);
    const bool vec = ((uintptr_t)date | (uintptr_t)qty | (uintptr_t)price | (uintptr_t)disc) % 16 == 0;
    if (vec)
        launch(c, "q6", k_q6<true>, dim3(nb), dim3(Q6_BLOCK), 0, s, date, qty, price, disc, n, a, part.as<uint64_t>());
    else
        launch(c, "q6", k_q6<false>, dim3(nb), dim3(Q6_BLOCK), 0, s, date, qty, price, disc, n, a, part.as<uint64_t>());
    launch(c, "q6_final", k_q6_final, dim3(1), dim3(256), 0, s, (const uint64_t*)part.as<uint64_t>(), nb,
           res.as<uint64_t>());
    uint64_t h[2];
    read_small(c, h, res.get(), 16, s);
    *revenue = (int64_t)h[0];
    return h[1];
}

uint64_t filter_groupby_q1(pimdal_ctx* c, const int32_t* date, const uint8_t* f1, const uint8_t* f2,
                           const int32_t* key, const int64_t* qty, const int64_t* price, const int64_t* disc,
                           const int64_t* tax, uint64_t n, int32_t cutoff, const pimdal_q1_result& out,
                           uint64_t cap, hipStream_t s) {
    PD_REQUIRE(date && qty && price && disc && tax && (key || (f1 && f2)), PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(out.key && out.sum_qty && out.sum_price && out.sum_disc && out.sum_disc_price && out.sum_charge &&
                   out.count, PIMDAL_E_INVALID, "missing output column");
    if (n == 0) return 0;
    if (!key) {
        // dense path: flag codes from a 64K-row sample, then filter + Q1
        // measures in register accumulators for up to 8 codes; a kept row
        // whose flag byte the sample lacks sets `miss`, and the call is then
        // redone with the codes of every row.  One readback at the end.
        // scratch: ng u64 | miss | reduction counter | meta[2] | bits[16] | lut[512] | inv[512]
        Scratch small(c, 8 + 8 + 8 + 64 + 512 + 512, s);
        uint64_t* ng = small.as<uint64_t>();
        uint32_t* miss = reinterpret_cast<uint32_t*>(ng + 1);
        uint32_t* done_ctr = miss + 1;
        uint32_t* meta = miss + 2;
        uint32_t* bits = meta + 2;
        uint8_t* lut = reinterpret_cast<uint8_t*>(bits + 16);
        uint8_t* inv = lut + 512;
        const uint32_t nb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, QD_BLOCK * 16), (uint64_t)c->cu_count * 8));
        Scratch part(c, (uint64_t)nb * 8 * 8 * 8, s), sums(c, 8 * 6 * 16, s);
        const bool vec = ((uintptr_t)date | (uintptr_t)f1 | (uintptr_t)f2 | (uintptr_t)qty | (uintptr_t)price |
                          (uintptr_t)disc | (uintptr_t)tax) % 16 == 0;
        // codes of the first `rows` rows; the sample's kernel also clears the
        // miss flag and the reduction counter, the full pass clears them apart
        auto codes = [&](uint64_t rows) {
            if (rows <= 65536) {
                launch(c, "q1_lut", k_flag_sample_codes, dim3(1), dim3(1024), 0, s, f1, f2, (uint32_t)rows, lut, inv,
                       meta, miss);
                return;
            }
            fill_u32(c, miss, 0, 4 + 16, s);   // miss, counter, meta[2] (rewritten below), bits[16]
            const uint32_t gd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(rows, 256 * 64), (uint64_t)c->cu_count * 4));
            launch(c, "q1_domain", k_flag_domain, dim3(gd), dim3(256), 0, s, f1, f2, rows, bits);
            launch(c, "q1_lut", k_flag_lut, dim3(1), dim3(64), 0, s, (const uint32_t*)bits, lut, inv, meta);
        };
        auto dense = [&](auto gtag) {
            constexpr int GG = decltype(gtag)::value;
            if (vec)
                launch(c, "q1_dense", k_q1_dense<GG, true>, dim3(nb), dim3(QD_BLOCK), 0, s, date, f1, f2, qty, price, disc,
                       tax, n, cutoff, (const uint8_t*)lut, (const uint32_t*)meta, part.as<uint64_t>(), miss);
            else
                launch(c, "q1_dense", k_q1_dense<GG, false>, dim3(nb), dim3(QD_BLOCK), 0, s, date, f1, f2, qty, price,
                       disc, tax, n, cutoff, (const uint8_t*)lut, (const uint32_t*)meta, part.as<uint64_t>(), miss);
            launch(c, "q1_reduce", k_q1_reduce, dim3(GG), dim3(256), 0, s, (const uint64_t*)part.as<uint64_t>(), nb,
                   (uint32_t)GG, sums.as<unsigned __int128>(), done_ctr, (const uint32_t*)meta, (const uint8_t*)inv,
                   out.key, out.sum_qty, out.sum_price, out.sum_disc, out.sum_disc_price, out.sum_charge, out.count, cap,
                   ng);
        };
        struct { uint64_t ng; uint32_t miss, pad, n1, n2; } hv;
        auto readback = [&]() {
            read_small(c, &hv, ng, sizeof(hv), s);
        };
        codes(std::min<uint64_t>(n, 65536));
        dense(std::integral_constant<int, 8>{});
        readback();
        bool done = hv.n1 * hv.n2 <= 8 && !hv.miss;
        if (!done && hv.n1 * hv.n2 <= 8) {
            // the sample missed a flag value: codes of every row
            codes(n);
            read_small(c, &hv.n1, meta, 8, s);
            const uint32_t groups = hv.n1 * hv.n2;
            if (groups <= 8) {
                // accumulators for exactly as many codes as the data has
                // (rounded up to even): fewer selects per row
                if (groups <= 2) dense(std::integral_constant<int, 2>{});
                else if (groups <= 4) dense(std::integral_constant<int, 4>{});
                else if (groups <= 6) dense(std::integral_constant<int, 6>{});
                else dense(std::integral_constant<int, 8>{});
                readback();
                done = true;
            }
        }
        if (done) {
            PD_REQUIRE(hv.ng <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
            return hv.ng;
        }
    }
    GbSource src{};
    src.key = reinterpret_cast<const uint32_t*>(key);
    src.f1 = f1;
    src.f2 = f2;
    src.date = date;
    src.cutoff = cutoff;
    src.in[0] = qty; src.in[1] = price; src.in[2] = disc; src.in[3] = tax;
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "group-by supports up to 2^30-1 rows per call");
    if (key && c->tune_get("gb_packed", 1)) {
        // filter + pack to (key, 8-byte measures) in place of the rows, then
        // onesweep partitions and per-partition LDS tables writing the Q1
        // arrays directly; measures beyond the packed widths take the
        // general path below
        Scratch ovf(c, 8, s);
        fill_u32(c, ovf.get(), 0, 1, s);
        // The overflow flag is read after the packed path has run (its group
        // count depends on the keys only); when set, the general path below
        // recomputes every output.
        const OutQ1 o{out.key, out.sum_qty, out.sum_price, out.sum_disc, out.sum_disc_price, out.sum_charge, out.count};
        // LDS table per partition: 1024 slots (64 KB, two blocks per CU;
        // default) or 2048 slots (128 KB, one block per CU, half the
        // partitions)
        const uint64_t g =
            c->tune_get("gb_q1_slots", 1024) == 1024
                ? q1_key_packed<AggQ1T<1024>>(c, date, reinterpret_cast<const uint32_t*>(key), qty, price, disc, tax,
                                              n, cutoff, o, cap, ovf.as<uint32_t>(), s)
                : q1_key_packed<AggQ1T<2048>>(c, date, reinterpret_cast<const uint32_t*>(key), qty, price, disc, tax,
                                              n, cutoff, o, cap, ovf.as<uint32_t>(), s);
        uint32_t overflow = 0;
        read_small(c, &overflow, ovf.get(), 4, s);
        if (!overflow) return g;
    }
    const uint64_t stage = std::min<uint64_t>(n, std::max<uint64_t>(cap, 1));
    Scratch k32(c, stage * 4, s), fs(c, stage * 8 * AggQ1::NF, s);
    const uint64_t g = group_by<AggQ1>(c, src, n, k32.as<uint32_t>(), fs.as<uint64_t>(), stage, s);
    PD_REQUIRE(g <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    launch(c, "q1_unpack", k_q1_unpack, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(g, 256), 4096))),
           dim3(256), 0, s, g, (const uint32_t*)k32.as<uint32_t>(), (const uint64_t*)fs.as<uint64_t>(), stage,
           out.key, out.sum_qty, out.sum_price, out.sum_disc, out.sum_disc_price, out.sum_charge, out.count);
    PD_HIP(hipStreamSynchronize(s));
    return g;
}

// ---- asynchronous Q1 (no host reads: capturable into a plan) ----

// Result of the asynchronous Q1 forms, from device words: *d_ng = groups,
// *d_status = 0, PIMDAL_E_UNSUPPORTED when the data left what the form
// covers (dense: more than 8 flag combinations; by key: a measure beyond the
// packed widths, f0, or a partition beyond its LDS table, f1), or
// PIMDAL_E_CAPACITY when there are more groups than `cap` (only `cap`
// written).
__global__ void k_q1_async_status(const uint32_t* __restrict__ dense_meta, const uint32_t* __restrict__ f0,
                                  const uint32_t* __restrict__ f1, const uint64_t* __restrict__ groups, uint64_t cap,
                                  uint64_t* __restrict__ d_ng, int32_t* __restrict__ d_status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bool unsupported = false;
    if (dense_meta) unsupported |= dense_meta[0] * dense_meta[1] > 8u;
    if (f0) unsupported |= *f0 != 0;
    if (f1) unsupported |= *f1 != 0;
    const uint64_t g = groups ? *groups : 0;
    *d_ng = unsupported ? 0 : g;
    *d_status = unsupported ? PIMDAL_E_UNSUPPORTED : g > cap ? PIMDAL_E_CAPACITY : PIMDAL_OK;
}

// k_q1_unpack with the group count read on the device: the first
// min(*count, cap) staged groups.
__global__ void k_q1_unpack_dev(const uint64_t* __restrict__ count, uint64_t cap, const uint32_t* __restrict__ key,
                                const uint64_t* __restrict__ f, uint64_t stride, int32_t* __restrict__ okey,
                                int64_t* __restrict__ sq, int64_t* __restrict__ sp, int64_t* __restrict__ sd,
                                int64_t* __restrict__ sdp, int64_t* __restrict__ sch, int64_t* __restrict__ cnt) {
    const uint64_t g = min(*count, cap);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g; i += (uint64_t)gridDim.x * blockDim.x) {
        okey[i] = (int32_t)key[i];
        sq[i] = (int64_t)f[i];
        sp[i] = (int64_t)f[stride + i];
        sd[i] = (int64_t)f[2 * stride + i];
        sdp[i] = (int64_t)f[3 * stride + i];
        sch[2 * i] = (int64_t)f[4 * stride + i];
        sch[2 * i + 1] = (int64_t)f[5 * stride + i];
        cnt[i] = (int64_t)f[6 * stride + i];
    }
}

// Q1 by key without host reads: pack (overflow flag on the device),
// partition by hash digits (ordinary launches while capturing), one LDS
// table per partition.  Partitions beyond an LDS table and measures beyond
// the packed widths are reported through the status word instead of taking
// the general path.
template <typename Agg>
static void q1_key_async(pimdal_ctx* c, const int32_t* date, const int32_t* key, const int64_t* qty,
                         const int64_t* price, const int64_t* disc, const int64_t* tax, uint64_t n, int32_t cutoff,
                         const OutQ1& o, uint64_t cap, uint64_t* d_ng, int32_t* d_status, hipStream_t s) {
    Scratch ovf(c, 8, s);
    fill_u32(c, ovf.get(), 0, 2, s);
    const uint32_t bits = gb_part_bits(c, n, (int64_t)Agg::TS * 2);
    const uint32_t F = 1u << bits;
    Scratch qk(c, n * 4, s), qv(c, n * 8, s), bounds(c, ((uint64_t)F + 1) * 4, s);
    q1_pack_rows(c, date, reinterpret_cast<const uint32_t*>(key), qty, price, disc, tax, n, cutoff, bits,
                 qk.as<uint32_t>(), qv.as<uint64_t>(), bounds.as<uint32_t>(), ovf.as<uint32_t>(), s);
    Scratch meta(c, 8 * 4, s), tog(c, (uint64_t)F * 4, s);
    fill_u64(c, meta.get(), 0, 4, s);
    fill_u32(c, tog.get(), 0, F, s);
    const uint32_t max_rows = (uint32_t)c->tune_get("gb_lds_max_rows", 65536);
    const RowsQ1Packed src{qk.as<uint32_t>(), (const uint64_t*)qv.get()};
    const bool direct_out = cap >= n;
    Scratch sk(c, direct_out ? 0 : n * 4, s), sf(c, direct_out ? 0 : n * 8 * Agg::NF, s);
    if (direct_out) {
        launch(c, "gb_lds", k_gb_lds<Agg, RowsQ1Packed, OutQ1>, dim3(F), dim3(GB_LDS_BLOCK), 0, s, src,
               (const uint32_t*)bounds.as<uint32_t>(), 1u, F, max_rows, tog.as<uint32_t>(), o, meta.as<uint64_t>());
    } else {
        const OutFields stage{sk.as<uint32_t>(), sf.as<uint64_t>(), n};
        launch(c, "gb_lds", k_gb_lds<Agg, RowsQ1Packed, OutFields>, dim3(F), dim3(GB_LDS_BLOCK), 0, s, src,
               (const uint32_t*)bounds.as<uint32_t>(), 1u, F, max_rows, tog.as<uint32_t>(), stage, meta.as<uint64_t>());
        launch(c, "q1_unpack", k_q1_unpack_dev, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(cap, 256), 4096))),
               dim3(256), 0, s, (const uint64_t*)meta.as<uint64_t>(), cap, (const uint32_t*)sk.as<uint32_t>(),
               (const uint64_t*)sf.as<uint64_t>(), n, o.key, o.sq, o.sp, o.sd, o.sdp, o.sch, o.cnt);
    }
    launch(c, "q1_status", k_q1_async_status, dim3(1), dim3(64), 0, s, (const uint32_t*)nullptr,
           (const uint32_t*)ovf.as<uint32_t>(), (const uint32_t*)(meta.as<uint64_t>() + 1),
           (const uint64_t*)meta.as<uint64_t>(), cap, d_ng, d_status);
}

// The asynchronous Q1 forms.  Flags (key == NULL): the codes of every row's
// flag bytes (one pass over the two flag columns), then the dense pass for
// up to 8 codes and the reduction, which writes the groups in ascending key
// order; by key: q1_key_async.  Results and status come back in device
// words (k_q1_async_status).
void filter_groupby_q1_async(pimdal_ctx* c, const int32_t* date, const uint8_t* f1, const uint8_t* f2,
                             const int32_t* key, const int64_t* qty, const int64_t* price, const int64_t* disc,
                             const int64_t* tax, uint64_t n, int32_t cutoff, const pimdal_q1_result& out, uint64_t cap,
                             uint64_t* d_ng, int32_t* d_status, hipStream_t s) {
    PD_REQUIRE(date && qty && price && disc && tax && (key || (f1 && f2)), PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(out.key && out.sum_qty && out.sum_price && out.sum_disc && out.sum_disc_price && out.sum_charge &&
                   out.count, PIMDAL_E_INVALID, "missing output column");
    PD_REQUIRE(d_ng && d_status, PIMDAL_E_INVALID, "missing result word");
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "group-by supports up to 2^30-1 rows per call");
    if (n == 0) {
        launch(c, "q1_status", k_q1_async_status, dim3(1), dim3(64), 0, s, (const uint32_t*)nullptr,
               (const uint32_t*)nullptr, (const uint32_t*)nullptr, (const uint64_t*)nullptr, cap, d_ng, d_status);
        return;
    }
    if (key) {
        const OutQ1 o{out.key, out.sum_qty, out.sum_price, out.sum_disc, out.sum_disc_price, out.sum_charge, out.count};
        if (c->tune_get("gb_q1_slots", 1024) == 1024)
            q1_key_async<AggQ1T<1024>>(c, date, key, qty, price, disc, tax, n, cutoff, o, cap, d_ng, d_status, s);
        else
            q1_key_async<AggQ1T<2048>>(c, date, key, qty, price, disc, tax, n, cutoff, o, cap, d_ng, d_status, s);
        return;
    }
    // scratch: ng u64 | miss | reduction counter | meta[2] | bits[16] | lut[512] | inv[512]
    Scratch small(c, 8 + 8 + 8 + 64 + 512 + 512, s);
    uint64_t* ng = small.as<uint64_t>();
    uint32_t* miss = reinterpret_cast<uint32_t*>(ng + 1);
    uint32_t* done_ctr = miss + 1;
    uint32_t* meta = miss + 2;
    uint32_t* bits = meta + 2;
    uint8_t* lut = reinterpret_cast<uint8_t*>(bits + 16);
    uint8_t* inv = lut + 512;
    constexpr int GG = 8;
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, QD_BLOCK * 16), (uint64_t)c->cu_count * 8));
    Scratch part(c, (uint64_t)nb * 8 * 8 * 8, s), sums(c, 8 * 6 * 16, s);
    const bool vec = ((uintptr_t)date | (uintptr_t)f1 | (uintptr_t)f2 | (uintptr_t)qty | (uintptr_t)price |
                      (uintptr_t)disc | (uintptr_t)tax) % 16 == 0;
    fill_u32(c, miss, 0, 4 + 16, s);   // miss, counter, meta[2], bits[16]
    const uint32_t gd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 64), (uint64_t)c->cu_count * 4));
    launch(c, "q1_domain", k_flag_domain, dim3(gd), dim3(256), 0, s, f1, f2, n, bits);
    launch(c, "q1_lut", k_flag_lut, dim3(1), dim3(64), 0, s, (const uint32_t*)bits, lut, inv, meta);
    if (vec)
        launch(c, "q1_dense", k_q1_dense<GG, true>, dim3(nb), dim3(QD_BLOCK), 0, s, date, f1, f2, qty, price, disc, tax, n,
               cutoff, (const uint8_t*)lut, (const uint32_t*)meta, part.as<uint64_t>(), miss);
    else
        launch(c, "q1_dense", k_q1_dense<GG, false>, dim3(nb), dim3(QD_BLOCK), 0, s, date, f1, f2, qty, price, disc, tax,
               n, cutoff, (const uint8_t*)lut, (const uint32_t*)meta, part.as<uint64_t>(), miss);
    launch(c, "q1_reduce", k_q1_reduce, dim3(GG), dim3(256), 0, s, (const uint64_t*)part.as<uint64_t>(), nb, (uint32_t)GG,
           sums.as<unsigned __int128>(), done_ctr, (const uint32_t*)meta, (const uint8_t*)inv, out.key, out.sum_qty,
           out.sum_price, out.sum_disc, out.sum_disc_price, out.sum_charge, out.count, cap, ng);
    launch(c, "q1_status", k_q1_async_status, dim3(1), dim3(64), 0, s, (const uint32_t*)meta, (const uint32_t*)nullptr,
           (const uint32_t*)nullptr, (const uint64_t*)ng, cap, d_ng, d_status);
}

// Q6 with its results left in device words (revenue, rows); no host read.
void q6_async(pimdal_ctx* c, const int32_t* date, const int64_t* qty, const int64_t* price, const int64_t* disc,
              uint64_t n, const Q6Args& a, int64_t* d_revenue, uint64_t* d_rows, hipStream_t s) {
    PD_REQUIRE((date && qty && price && disc) || n == 0, PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(d_revenue && d_rows, PIMDAL_E_INVALID, "missing result word");
    const uint32_t nb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, Q6_BLOCK * 4 * 4),
                                                                           (uint64_t)c->cu_count * 8));
    Scratch part(c, (uint64_t)nb * 16, s), res(c, 16, s);
    const bool vec = ((uintptr_t)date | (uintptr_t)qty | (uintptr_t)price | (uintptr_t)disc) % 16 == 0;
    if (vec)
        launch(c, "q6", k_q6<true>, dim3(nb), dim3(Q6_BLOCK), 0, s, date, qty, price, disc, n, a, part.as<uint64_t>());
    else
        launch(c, "q6", k_q6<false>, dim3(nb), dim3(Q6_BLOCK), 0, s, date, qty, price, disc, n, a, part.as<uint64_t>());
    launch(c, "q6_final", k_q6_final, dim3(1), dim3(256), 0, s, (const uint64_t*)part.as<uint64_t>(), nb,
           res.as<uint64_t>());
    copy_d2d(c, d_revenue, res.get(), 8, s);
    copy_d2d(c, d_rows, res.as<uint64_t>() + 1, 8, s);
}

void flag_codes(pimdal_ctx* c, const uint8_t* f1, const uint8_t* f2, uint64_t n, uint8_t* lut_dev, uint32_t meta[2],
                uint8_t* inv_host, hipStream_t s) {
    Scratch small(c, 64 + 8, s);
    uint32_t* bits = small.as<uint32_t>();
    fill_u32(c, bits, 0, 18, s);
    const uint32_t gd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 64), (uint64_t)c->cu_count * 4));
    if (n) launch(c, "q1_domain", k_flag_domain, dim3(gd), dim3(256), 0, s, f1, f2, n, bits);
    launch(c, "q1_lut", k_flag_lut, dim3(1), dim3(64), 0, s, (const uint32_t*)bits, lut_dev, lut_dev + 512, bits + 16);
    PD_HIP(hipMemcpyAsync(meta, bits + 16, 8, hipMemcpyDeviceToHost, s));
    if (inv_host) read_small(c, inv_host, lut_dev + 512, 512, s);
}

uint64_t hash_aggr(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                   int64_t* out_sum, int64_t* out_cnt, uint64_t cap, hipStream_t s) {
    PD_REQUIRE((key && val) || n == 0, PIMDAL_E_INVALID, "missing input column");
    if (n == 0) return 0;
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "hash aggregation supports up to 2^30-1 rows per call");
    // keys within 8192 values: key-indexed LDS arrays
    uint64_t g = 0;
    if (aggr_sum_direct(c, key, val, n, out_key, out_sum, out_cnt, cap, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    // few groups: per-block LDS tables merged in one HBM table
    const int64_t hint = c->tune_get("gb_groups_hint", 0);
    if (c->tune_get("aggr_small", 1) && hint <= 4096 &&
        aggr_sum_small(c, key, val, n, out_key, out_sum, out_cnt, cap, false, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    // keys within 2^24 of each other: one range-partition pass
    if (aggr_sum_dense(c, key, val, n, out_key, out_sum, out_cnt, cap, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    // otherwise onesweep partitions of (key, value) by hash digits and one
    // LDS table per partition, writing the outputs directly
    g = group_by_onesweep<AggSum, RowsSum>(
        c, key, val, 8, n, [](uint32_t* k, void* v) { return RowsSum{k, (const int64_t*)v}; },
        OutSum{out_key, out_sum, out_cnt}, cap, s);
    PD_HIP(hipStreamSynchronize(s));
    return g;
}

// Sort aggregation: groups in ascending key order.  Keys within 8192
// values sum into key-indexed LDS arrays; with few groups the hash tables'
// groups are ordered by key (sorting the rows would only add passes); with
// keys in a range of at most 2^24, range partitions summed into key-indexed
// LDS arrays; otherwise the rows are radix sorted by key and reduced in one
// pass over the sorted runs.
uint64_t sort_aggr(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                   int64_t* out_sum, int64_t* out_cnt, uint64_t cap, hipStream_t s) {
    PD_REQUIRE((key && val) || n == 0, PIMDAL_E_INVALID, "missing input column");
    if (n == 0) return 0;
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "sort aggregation supports up to 2^30-1 rows per call");
    uint64_t g = 0;
    if (aggr_sum_direct(c, key, val, n, out_key, out_sum, out_cnt, cap, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    const int64_t hint = c->tune_get("gb_groups_hint", 0);
    if (c->tune_get("aggr_small", 1) && hint <= 4096 &&
        aggr_sum_small(c, key, val, n, out_key, out_sum, out_cnt, cap, true, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    // keys within 2^24 of each other: range partitions give key order
    if (aggr_sum_dense(c, key, val, n, out_key, out_sum, out_cnt, cap, &g, s)) {
        PD_HIP(hipStreamSynchronize(s));
        return g;
    }
    Scratch sk(c, n * 4, s), sv(c, n * 8, s);
    sort_pairs(c, key, val, 8, n, sk.as<uint32_t>(), sv.get(), -1, -1, s);
    g = aggr_sum_sorted(c, sk.as<uint32_t>(), sv.as<int64_t>(), n, out_key, out_sum, out_cnt, cap, s);
    PD_HIP(hipStreamSynchronize(s));
    return g;
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_hash_aggr_dev(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n,
                                    uint32_t* out_key, int64_t* out_sum, int64_t* out_count, uint64_t capacity,
                                    uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = hash_aggr(c, key, val, n, out_key, out_sum, out_count, capacity, c->stream);
    PD_API_END
}

extern "C" int pimdal_sort_aggr_dev(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n,
                                    uint32_t* out_key, int64_t* out_sum, int64_t* out_count, uint64_t capacity,
                                    uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = sort_aggr(c, key, val, n, out_key, out_sum, out_count, capacity, c->stream);
    PD_API_END
}

extern "C" int pimdal_filter_groupby_q1_dev(pimdal_ctx* c, const int32_t* date, const uint8_t* flag1,
                                            const uint8_t* flag2, const int32_t* key, const int64_t* quantity,
                                            const int64_t* price, const int64_t* discount, const int64_t* tax,
                                            uint64_t n, int32_t cutoff, const pimdal_q1_result* out,
                                            uint64_t capacity, uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && out && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = filter_groupby_q1(c, date, flag1, flag2, key, quantity, price, discount, tax, n, cutoff, *out,
                                  capacity, c->stream);
    PD_API_END
}

extern "C" int pimdal_q6_dev(pimdal_ctx* c, const int32_t* date, const int64_t* quantity, const int64_t* price,
                             const int64_t* discount, uint64_t n, int32_t date_lo, int32_t date_hi, int64_t disc_lo,
                             int64_t disc_hi, int64_t qty_lt, int64_t* revenue, uint64_t* n_rows) {
    PD_API_BEGIN
    PD_REQUIRE(c && revenue && n_rows, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    const Q6Args a{date_lo, date_hi, disc_lo, disc_hi, qty_lt};
    *n_rows = q6(c, date, quantity, price, discount, n, a, revenue, c->stream);
    PD_API_END
}

extern "C" int pimdal_filter_groupby_q1_dev_async(pimdal_ctx* c, const int32_t* date, const uint8_t* flag1,
                                                  const uint8_t* flag2, const int32_t* key, const int64_t* quantity,
                                                  const int64_t* price, const int64_t* discount, const int64_t* tax,
                                                  uint64_t n, int32_t cutoff, const pimdal_q1_result* out,
                                                  uint64_t capacity, uint64_t* d_n_groups, int32_t* d_status) {
    PD_API_BEGIN
    PD_REQUIRE(c && out, PIMDAL_E_INVALID, "NULL argument");
    filter_groupby_q1_async(c, date, flag1, flag2, key, quantity, price, discount, tax, n, cutoff, *out, capacity,
                            d_n_groups, d_status, c->stream);
    PD_API_END
}

extern "C" int pimdal_q6_dev_async(pimdal_ctx* c, const int32_t* date, const int64_t* quantity, const int64_t* price,
                                   const int64_t* discount, uint64_t n, int32_t date_lo, int32_t date_hi,
                                   int64_t disc_lo, int64_t disc_hi, int64_t qty_lt, int64_t* d_revenue,
                                   uint64_t* d_rows) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "NULL argument");
    const Q6Args a{date_lo, date_hi, disc_lo, disc_hi, qty_lt};
    q6_async(c, date, quantity, price, discount, n, a, d_revenue, d_rows, c->stream);
    PD_API_END
}

// ---------------- column-level aggregation ----------------

namespace {

using AggFn = uint64_t (*)(pimdal_ctx*, const uint32_t*, const int64_t*, uint64_t, uint32_t*, int64_t*, int64_t*,
                           uint64_t, hipStream_t);

int aggr_cols(AggFn fn, pimdal_ctx* c, const pimdal_col* key, const pimdal_col* val, pimdal_col** ok,
              pimdal_col** os, pimdal_col** oc) {
    PD_API_BEGIN
    PD_REQUIRE(c && key && val && ok && os && oc, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    PD_REQUIRE(type_width(key->type) == 4, PIMDAL_E_UNSUPPORTED, "group key must be 32-bit");
    PD_REQUIRE(val->type == PIMDAL_I64 || val->type == PIMDAL_U64, PIMDAL_E_UNSUPPORTED, "values must be 64-bit");
    PD_REQUIRE(key->length == val->length, PIMDAL_E_INVALID, "column lengths differ");
    if (key->validity || val->validity) {
        // NULL keys form one group (last); SUM skips NULL values, COUNT counts
        // the non-NULL ones; groups in ascending key order
        aggr_cols_nullable(c, key, val, ok, os, oc);
        return PIMDAL_OK;
    }
    const uint64_t n = key->length, cap = std::max<uint64_t>(n, 1);
    pimdal_col *a = nullptr, *b = nullptr, *d = nullptr;
    try {
        a = new pimdal_col{key->type, 0, cap, c->pool.alloc(cap * 4, c->stream), true};
        b = new pimdal_col{PIMDAL_I64, 0, cap, c->pool.alloc(cap * 8, c->stream), true};
        d = new pimdal_col{PIMDAL_I64, 0, cap, c->pool.alloc(cap * 8, c->stream), true};
        const uint64_t g = fn(c, (const uint32_t*)key->data, (const int64_t*)val->data, n, (uint32_t*)a->data,
                              (int64_t*)b->data, (int64_t*)d->data, cap, c->stream);
        a->length = b->length = d->length = g;
    } catch (...) {
        for (pimdal_col* x : {a, b, d})
            if (x) { c->pool.free(x->data, c->stream); delete x; }
        throw;
    }
    *ok = a;
    *os = b;
    *oc = d;
    PD_API_END
}

}  // namespace

extern "C" int pimdal_hash_aggr(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* val, pimdal_col** ok,
                                pimdal_col** os, pimdal_col** oc) {
    return aggr_cols(pd::hash_aggr, c, key, val, ok, os, oc);
}

extern "C" int pimdal_sort_aggr(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* val, pimdal_col** ok,
                                pimdal_col** os, pimdal_col** oc) {
    return aggr_cols(pd::sort_aggr, c, key, val, ok, os, oc);
}
