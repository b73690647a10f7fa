// q3.hip — the TPC-H Q3 shape: two joins chained on the device with the
// group-by fused into the second join's probe; further down the TPC-H Q4
// shape (a semi-join of orders with lineitems, counted per priority) and
// the TPC-H Q5 shape (five tables joined, revenue per nation).
//
// Reference semantics: pimdal/main/tpc_h/query3/host/host_q3.cpp:471-519
// (five DPU programs launched in turn, the host collecting and
// redistributing rows between them): kernel_q3_1.c:116-137 selects the
// customers of one market segment; kernel_q3_2.c:146-190 the orders with
// o_orderdate < DATE; kernel_q3_3.c:160-215 hash-joins them on custkey;
// kernel_q3_4.c:120-141 selects the lineitems with l_shipdate > DATE;
// kernel_q3_5.c:66-120, 170-262 hash-joins those with the orders on
// orderkey, computes revenue = l_extendedprice * (100 - l_discount), sums it
// per (orderkey, orderdate, shippriority) and sorts for the top 10
// (host_q3.cpp:300-345 merges the devices' top rows).
//
// Plan on gfx950 (nothing leaves the device between the steps):
//  1. k_q3_cust    the qualifying customers' keys: those below 2^25 as
//                  bits of a 4 MB bitmap (an L2 read per probe for dense
//                  custkeys), larger ones into a hash set in HBM;
//  2. k_q3_osel / k_q3_orders   orders with o_orderdate < date whose
//                  customer is in the set: listed (one pass over the orders),
//                  then inserted into an open-addressing table (orderkey,
//                  order row) sized for exactly that count;
//  3. k_q3_lineitem   one streaming pass over the lineitems: shipdate >
//                  date, probe the orders table, and add the revenue into
//                  the order's slot (the join and the group-by in one pass:
//                  the group key (orderkey, orderdate, shippriority) is the
//                  order row, unique per orderkey);
//  4. k_q3_emit    the slots that received a lineitem, compacted, with the
//                  order's date and ship priority gathered;
//  5. the TPC-H order (revenue descending, o_orderdate, then orderkey
//     ascending) by four stable radix sorts of a row permutation, least
//     significant key first; the first `capacity` groups are written.
#include <algorithm>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

__device__ __forceinline__ uint64_t q3_slot(uint32_t key, uint64_t mask) { return pd_mix64_dev(key) & mask; }

// Customer set: keys below 2^Q3_CBITS as bits of a 4 MB bitmap (TPC-H's
// dense custkeys touch 1.5e6 bits at SF10, which stay in L2, where a probe
// of the 32 MB hash set is a random MALL read), larger keys in the hash set
// (slot word = 1 << 32 | custkey, 0 = empty).
constexpr uint32_t Q3_CBITS = 25;   // the largest bitmap: 2^25 keys, 4 MB

__global__ void k_q3_cust(const int32_t* __restrict__ custkey, const uint8_t* __restrict__ segment, uint64_t n,
                          uint32_t code, uint32_t* __restrict__ cbits, uint32_t dn, uint64_t* __restrict__ set,
                          uint64_t mask) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (segment[i] != code) continue;
        const uint32_t k = (uint32_t)custkey[i];
        if (k < dn) {
            atomicOr(&cbits[k >> 5], 1u << (k & 31));
            continue;
        }
        const uint64_t want = (1ull << 32) | k;
        uint64_t h = q3_slot(k, mask);
        while (true) {
            const uint64_t prev = atomicCAS((unsigned long long*)&set[h], 0ull, (unsigned long long)want);
            if (prev == 0 || prev == want) break;
            h = (h + 1) & mask;
        }
    }
}

// Qualifying orders (o_orderdate < date, customer in the set) as a list of
// order rows.  Block b takes the Q3_CHUNK rows from b * Q3_CHUNK, each
// thread Q3_CHUNK / 256 of them (consecutive groups of 4, all probes of a
// group in flight together); one counter atomic per block claims the
// block's output range (same-word atomics serialize, so not one per wave).
constexpr uint32_t Q3_CHUNK = 8192;

__global__ __launch_bounds__(256) void k_q3_osel(const int32_t* __restrict__ odate, const int32_t* __restrict__ ocust,
                                                 uint64_t n, int32_t date, const uint32_t* __restrict__ cbits,
                                                 uint32_t dn, const uint64_t* __restrict__ set, uint64_t mask,
                                                 unsigned long long* __restrict__ ctr, uint32_t* __restrict__ rows) {
    constexpr int PER = Q3_CHUNK / 256;   // 32 rows per thread
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_base;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t r0 = (uint64_t)blockIdx.x * Q3_CHUNK;
    if (r0 >= n) return;   // no rows (the loads below are clamped to row n - 1)
    uint32_t bits = 0;   // bit j: row r0 + j * 256 + t qualifies
#pragma unroll
    for (int j0 = 0; j0 < PER; j0 += 4) {
        bool cand[4];
        uint32_t key[4];
        uint64_t h[4], e[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // loads clamped to the last row, so none waits behind a branch
            const uint64_t i = r0 + (uint64_t)(j0 + q) * 256 + t;
            const uint64_t ci = i < n ? i : n - 1;
            const int32_t d = odate[ci];
            const uint32_t cu = (uint32_t)ocust[ci];
            cand[q] = i < n && d < date;
            key[q] = cand[q] ? cu : 0u;
            h[q] = q3_slot(key[q], mask);
        }
        bool small[4];
        uint32_t cw[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            small[q] = key[q] < dn;
            cw[q] = cand[q] && small[q] ? cbits[key[q] >> 5] : 0u;
            e[q] = cand[q] && !small[q] ? set[h[q]] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!cand[q]) continue;
            if (small[q]) {
                if ((cw[q] >> (key[q] & 31)) & 1u) bits |= 1u << (j0 + q);
                continue;
            }
            const uint64_t want = (1ull << 32) | key[q];
            while (e[q] != 0 && e[q] != want) {
                h[q] = (h[q] + 1) & mask;
                e[q] = set[h[q]];
            }
            if (e[q] == want) bits |= 1u << (j0 + q);
        }
    }
    const uint32_t cnt = __popc(bits);
    const uint32_t inc = pd::wave_incl_scan(cnt);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t before = inc - cnt, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        before += (i < (int)w) ? s_w[i] : 0u;
        tot += s_w[i];
    }
    if (t == 0) s_base = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0ull;
    __syncthreads();
    uint64_t o = s_base + before;
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1;
        rows[o++] = (uint32_t)(r0 + (uint64_t)j * 256 + t);
    }
}

// Orders table from the qualifying rows: slot word = (order row + 1) << 32
// | orderkey (0 = empty); a second orderkey equal to one already in the
// table sets *dup.
__global__ void k_q3_orders(const uint32_t* __restrict__ rows, uint64_t nq, const int32_t* __restrict__ okey,
                            uint64_t* __restrict__ otab, uint64_t omask, uint32_t* __restrict__ dup) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nq; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = rows[j];
        const uint32_t k = (uint32_t)okey[i];
        const uint64_t want = ((i + 1) << 32) | k;
        uint64_t h = q3_slot(k, omask);
        while (true) {
            const uint64_t prev = atomicCAS((unsigned long long*)&otab[h], 0ull, (unsigned long long)want);
            if (prev == 0) break;
            if ((uint32_t)prev == k) { atomicOr(dup, 1u); break; }
            h = (h + 1) & omask;
        }
    }
}

// One pass over the lineitems: revenue of every row with shipdate > date
// whose order is in the table, added to the order's slot; the slot's hit
// bit marks a group.  Four rows per thread, all probes in flight together.
__global__ __launch_bounds__(256) void k_q3_lineitem(const int32_t* __restrict__ lkey,
                                                     const int32_t* __restrict__ lship,
                                                     const int64_t* __restrict__ lprice,
                                                     const int64_t* __restrict__ ldisc, uint64_t n, int32_t date,
                                                     const uint64_t* __restrict__ otab, uint64_t omask,
                                                     unsigned long long* __restrict__ rev, uint32_t* __restrict__ hit) {
    constexpr int R = 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; base < n; base += stride * R) {
        uint32_t key[R];
        bool want[R];
        uint64_t h[R], e[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            // loads clamped to the last row, so none waits behind a branch
            const uint64_t i = base + (uint64_t)r * stride;
            const uint64_t ci = i < n ? i : n - 1;
            const int32_t d = lship[ci];
            const uint32_t k = (uint32_t)lkey[ci];
            want[r] = i < n && d > date;
            key[r] = want[r] ? k : 0u;
            h[r] = q3_slot(key[r], omask);
        }
#pragma unroll
        for (int r = 0; r < R; r++) e[r] = want[r] ? otab[h[r]] : 0ull;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!want[r]) continue;
            while (e[r] != 0 && (uint32_t)e[r] != key[r]) {
                h[r] = (h[r] + 1) & omask;
                e[r] = otab[h[r]];
            }
            if (e[r] == 0) continue;
            const uint64_t i = base + (uint64_t)r * stride;
            const int64_t v = lprice[i] * (100 - ldisc[i]);
            atomicAdd(&rev[h[r]], (unsigned long long)v);
            const uint32_t bit = 1u << (h[r] & 31);
            if (!(hit[h[r] >> 5] & bit)) atomicOr(&hit[h[r] >> 5], bit);
        }
    }
}

// Groups: the hit slots, compacted, with the order's key, date and ship
// priority.  Block b takes the 8192 slots of hit words b * 256 .. + 255,
// thread t one word; a block scan of the words' popcounts and one counter
// atomic per block give every group its output position.
__global__ __launch_bounds__(256) void k_q3_emit(const uint64_t* __restrict__ otab, const int64_t* __restrict__ rev,
                                                 const uint32_t* __restrict__ hit, uint64_t words,
                                                 const int32_t* __restrict__ odate,
                                                 const int32_t* __restrict__ oprio,
                                                 unsigned long long* __restrict__ ctr, uint32_t* __restrict__ g_key,
                                                 int64_t* __restrict__ g_rev, int32_t* __restrict__ g_date,
                                                 int32_t* __restrict__ g_prio) {
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_base;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t wd = (uint64_t)blockIdx.x * 256 + t;
    uint32_t bits = wd < words ? hit[wd] : 0u;
    const uint32_t cnt = __popc(bits);
    const uint32_t inc = pd::wave_incl_scan(cnt);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t before = inc - cnt, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        before += (i < (int)w) ? s_w[i] : 0u;
        tot += s_w[i];
    }
    if (t == 0) s_base = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0ull;
    __syncthreads();
    uint64_t o = s_base + before;
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1;
        const uint64_t s = wd * 32 + (uint64_t)j;
        const uint64_t e = otab[s];
        const uint64_t row = (e >> 32) - 1;
        g_key[o] = (uint32_t)e;
        g_rev[o] = rev[s];
        g_date[o] = odate[row];
        g_prio[o] = oprio[row];
        o++;
    }
}

// Sort keys of the TPC-H order, gathered through the current permutation:
// 0 orderkey, 1 orderdate (sign bit flipped), 2 / 3 low / high word of
// ~(revenue with its sign bit flipped) (ascending = revenue descending).
__global__ void k_q3_sortkey(uint32_t which, const uint32_t* __restrict__ perm, uint64_t n,
                             const uint32_t* __restrict__ g_key, const int64_t* __restrict__ g_rev,
                             const int32_t* __restrict__ g_date, uint32_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        uint32_t k;
        if (which == 0) k = g_key[p];
        else if (which == 1) k = (uint32_t)g_date[p] ^ 0x80000000u;
        else {
            const uint64_t r = ~((uint64_t)g_rev[p] ^ (1ull << 63));
            k = which == 2 ? (uint32_t)r : (uint32_t)(r >> 32);
        }
        out[i] = k;
    }
}

__global__ void k_q3_out(const uint32_t* __restrict__ perm, uint64_t n, const uint32_t* __restrict__ g_key,
                         const int64_t* __restrict__ g_rev, const int32_t* __restrict__ g_date,
                         const int32_t* __restrict__ g_prio, int32_t* __restrict__ o_key, int64_t* __restrict__ o_rev,
                         int32_t* __restrict__ o_date, int32_t* __restrict__ o_prio) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[i];
        o_key[i] = (int32_t)g_key[p];
        o_rev[i] = g_rev[p];
        o_date[i] = g_date[p];
        o_prio[i] = g_prio[p];
    }
}

__global__ void k_q3_iota(uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)i;
}

// ---- the first `cap` groups when cap is small (the reference's top 10) ----
// Order key of a group: r = ~(revenue ^ 2^63) (ascending = revenue
// descending), then orderdate ^ 2^31, then orderkey.  A histogram of r over
// its range finds the bins holding the first cap groups; those candidates
// (every group in a bin at or below the bin of the cap-th one, ties
// included) are sorted whole in one block.
constexpr uint32_t Q3_TBINS = 4096;
constexpr uint32_t Q3_TCAND = 4096;   // candidates one block sorts

__device__ __forceinline__ uint64_t q3_rkey(int64_t rev) { return ~((uint64_t)rev ^ (1ull << 63)); }

// mm[0] = min r, mm[1] = max r (device atomics, one per block).
// Per-block (min, max) of the sort keys into part[2 b], part[2 b + 1]
// (one device atomic per wave on two words serialized the grid's waves:
// 0.045 ms at 122058 groups); k_q3_tfold folds them.
__device__ __forceinline__ void q3_block_minmax(uint64_t& lo, uint64_t& hi) {
    __shared__ uint64_t s_lo[256 / PD_WAVE], s_hi[256 / PD_WAVE];
    for (int o = PD_WAVE / 2; o > 0; o /= 2) {
        lo = min(lo, (uint64_t)__shfl_xor((unsigned long long)lo, o));
        hi = max(hi, (uint64_t)__shfl_xor((unsigned long long)hi, o));
    }
    const uint32_t w = threadIdx.x / PD_WAVE;
    if (pd::lane() == 0) {
        s_lo[w] = lo;
        s_hi[w] = hi;
    }
    __syncthreads();
    for (uint32_t i = 0; i < 256 / PD_WAVE; i++) {
        lo = min(lo, s_lo[i]);
        hi = max(hi, s_hi[i]);
    }
}

__global__ __launch_bounds__(256) void k_q3_tminmax(const int64_t* __restrict__ g_rev, uint64_t g,
                                                   unsigned long long* __restrict__ part) {
    uint64_t lo = ~0ull, hi = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = q3_rkey(g_rev[i]);
        lo = min(lo, r);
        hi = max(hi, r);
    }
    q3_block_minmax(lo, hi);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

// mm[0], mm[1] = the (min, max) of nb per-block partials.  One block.
__global__ __launch_bounds__(256) void k_q3_tfold(const unsigned long long* __restrict__ part, uint32_t nb,
                                                 unsigned long long* __restrict__ mm) {
    uint64_t lo = ~0ull, hi = 0;
    for (uint32_t b = threadIdx.x; b < nb; b += 256) {
        lo = min(lo, (uint64_t)part[2 * b]);
        hi = max(hi, (uint64_t)part[2 * b + 1]);
    }
    q3_block_minmax(lo, hi);
    if (threadIdx.x == 0) {
        mm[0] = lo;
        mm[1] = hi;
    }
}

__device__ __forceinline__ uint32_t q3_tshift(uint64_t span) {
    // smallest shift with span >> shift < Q3_TBINS
    uint32_t sh = 0;
    while ((span >> sh) >= Q3_TBINS) sh++;
    return sh;
}

__global__ __launch_bounds__(256) void k_q3_thist(const int64_t* __restrict__ g_rev, uint64_t g,
                                                 const unsigned long long* __restrict__ mm,
                                                 uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[Q3_TBINS];
    for (uint32_t i = threadIdx.x; i < Q3_TBINS; i += 256) h[i] = 0;
    __syncthreads();
    const uint64_t lo = mm[0];
    const uint32_t sh = q3_tshift(mm[1] - lo);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[(q3_rkey(g_rev[i]) - lo) >> sh], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < Q3_TBINS; i += 256)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// The bin b of the cap-th group and the groups in bins <= b: thr[0] = b,
// thr[1] = that count.  One block: thread t sums bins [16t, 16t + 16).
__global__ __launch_bounds__(256) void k_q3_tthresh(const uint32_t* __restrict__ hist, uint64_t cap,
                                                   uint32_t* __restrict__ thr) {
    __shared__ uint32_t s_w[4];
    constexpr uint32_t PER = Q3_TBINS / 256;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < PER; j++) sum += hist[t * PER + j];
    const uint32_t inc = pd::wave_incl_scan(sum);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (uint32_t i = 0; i < w; i++) run += s_w[i];
    // the thread whose bins reach cap finds the bin
    if (run < cap && run + sum >= cap) {
        for (uint32_t j = 0; j < PER; j++) {
            run += hist[t * PER + j];
            if (run >= cap) {
                thr[0] = t * PER + j;
                thr[1] = run;
                break;
            }
        }
    }
}

// Candidates: the groups whose bin is at most thr[0], listed (any order).
__global__ __launch_bounds__(256) void k_q3_tcand(const int64_t* __restrict__ g_rev, uint64_t g,
                                                 const unsigned long long* __restrict__ mm,
                                                 const uint32_t* __restrict__ thr, uint32_t* __restrict__ ctr,
                                                 uint32_t* __restrict__ cand) {
    const uint64_t lo = mm[0];
    const uint32_t sh = q3_tshift(mm[1] - lo), b = thr[0];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // whole waves step together, so the ballot and the wave's one counter
    // add see every lane
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x; base < g; base += stride) {
        const uint64_t i = base + threadIdx.x;
        const bool in = i < g && ((q3_rkey(g_rev[i]) - lo) >> sh) <= b;
        const uint64_t m = pd::ballot(in);
        uint32_t o = 0;
        if (pd::lane() == 0 && m) o = atomicAdd(ctr, (uint32_t)__popcll(m));
        o = __shfl(o, 0) + pd::popc64(m & pd::lanemask_lt());
        if (in) cand[o] = (uint32_t)i;
    }
}

// Sorts the nc candidates by (r, date, orderkey) in one block (bitonic over
// the next power of two, padding last) and writes the first k groups.
__global__ __launch_bounds__(1024) void k_q3_tsort(const uint32_t* __restrict__ cand, uint32_t nc, uint32_t k,
                                                  const uint32_t* __restrict__ g_key, const int64_t* __restrict__ g_rev,
                                                  const int32_t* __restrict__ g_date, const int32_t* __restrict__ g_prio,
                                                  int32_t* __restrict__ o_key, int64_t* __restrict__ o_rev,
                                                  int32_t* __restrict__ o_date, int32_t* __restrict__ o_prio) {
    __shared__ uint64_t s_r[Q3_TCAND];
    __shared__ uint64_t s_dk[Q3_TCAND];   // date ^ 2^31 << 32 | orderkey
    __shared__ uint32_t s_i[Q3_TCAND];
    uint32_t N = 2;
    while (N < nc) N <<= 1;
    for (uint32_t j = threadIdx.x; j < N; j += 1024) {
        if (j < nc) {
            const uint32_t i = cand[j];
            s_r[j] = q3_rkey(g_rev[i]);
            s_dk[j] = ((uint64_t)((uint32_t)g_date[i] ^ 0x80000000u) << 32) | g_key[i];
            s_i[j] = i;
        } else {
            s_r[j] = ~0ull;
            s_dk[j] = ~0ull;
            s_i[j] = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    for (uint32_t size = 2; size <= N; size <<= 1) {
        for (uint32_t stride = size / 2; stride > 0; stride >>= 1) {
            for (uint32_t j = threadIdx.x; j < N / 2; j += 1024) {
                const uint32_t a = 2 * j - (j & (stride - 1)), bb = a + stride;
                const bool up = (a & size) == 0;
                const bool gt = s_r[a] > s_r[bb] || (s_r[a] == s_r[bb] && s_dk[a] > s_dk[bb]);
                if (gt == up) {
                    const uint64_t r = s_r[a], dk = s_dk[a];
                    const uint32_t ii = s_i[a];
                    s_r[a] = s_r[bb]; s_dk[a] = s_dk[bb]; s_i[a] = s_i[bb];
                    s_r[bb] = r; s_dk[bb] = dk; s_i[bb] = ii;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t j = threadIdx.x; j < k; j += 1024) {
        const uint32_t i = s_i[j];
        o_key[j] = (int32_t)g_key[i];
        o_rev[j] = g_rev[i];
        o_date[j] = g_date[i];
        o_prio[j] = g_prio[i];
    }
}

static uint64_t q3_pow2_slots(uint64_t n) {
    uint64_t s = 1024;
    while (s < 2 * n) s <<= 1;
    return s;
}

// Keys a direct customer structure covers: the power of two at or above
// n_customer + 1 (TPC-H custkeys are 1 .. n_customer), from 2^16 up to
// 2^maxbits; larger keys go to the hash tables.
static uint32_t q3_direct_keys(uint64_t n_customer, uint32_t maxbits) {
    uint64_t d = 1ull << 16;
    while (d < n_customer + 1 && d < (1ull << maxbits)) d <<= 1;
    return (uint32_t)d;
}

static dim3 q3_grid(pimdal_ctx* c, uint64_t n) {
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), (uint64_t)c->cu_count * 16)));
}

uint64_t tpch_q3(pimdal_ctx* c, const pimdal_q3_tables& t, uint32_t segment, int32_t date, const pimdal_q3_result& out,
                 uint64_t cap, hipStream_t s) {
    PD_REQUIRE(t.c_custkey && t.c_mktsegment && t.o_orderkey && t.o_custkey && t.o_orderdate && t.o_shippriority &&
                   t.l_orderkey && t.l_extendedprice && t.l_discount && t.l_shipdate,
               PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(!cap || (out.orderkey && out.revenue && out.orderdate && out.shippriority), PIMDAL_E_INVALID,
               "missing output column");
    PD_REQUIRE(t.n_orders < (1ull << 30) && t.n_customer < (1ull << 31) && t.n_lineitem < (1ull << 40),
               PIMDAL_E_INVALID, "tables too large (at most 2^30-1 orders)");
    // 1. customer set
    const uint64_t cslots = q3_pow2_slots(t.n_customer);
    const uint32_t dn = q3_direct_keys(t.n_customer, Q3_CBITS);
    Scratch cset(c, cslots * 8, s), cbits(c, dn / 8, s), small(c, 64, s);
    fill_u64(c, cset.get(), 0, cslots, s);
    fill_u32(c, cbits.get(), 0, dn / 32, s);
    fill_u64(c, small.get(), 0, 8, s);
    unsigned long long* cnt = small.as<unsigned long long>();   // [0] qualifying orders, [1] groups
    uint32_t* dup = reinterpret_cast<uint32_t*>(small.as<uint64_t>() + 2);
    launch(c, "q3_cust", k_q3_cust, q3_grid(c, t.n_customer), dim3(256), 0, s, t.c_custkey, t.c_mktsegment,
           t.n_customer, segment, cbits.as<uint32_t>(), dn, cset.as<uint64_t>(), cslots - 1);
    // 2. qualifying orders: their rows listed, then a table sized for them
    Scratch orows(c, std::max<uint64_t>(t.n_orders, 1) * 4, s);
    launch(c, "q3_osel", k_q3_osel, dim3((uint32_t)std::max<uint64_t>(1, ceil_div(t.n_orders, Q3_CHUNK))), dim3(256), 0,
           s, t.o_orderdate, t.o_custkey, t.n_orders, date, (const uint32_t*)cbits.as<uint32_t>(), dn,
           (const uint64_t*)cset.as<uint64_t>(), cslots - 1, cnt, orows.as<uint32_t>());
    uint64_t nq = 0;
    read_small(c, &nq, cnt, 8, s);
    const uint64_t oslots = q3_pow2_slots(nq);
    Scratch otab(c, oslots * 8, s), rev(c, oslots * 8, s), hit(c, oslots / 8, s);
    fill_u64(c, otab.get(), 0, oslots, s);
    fill_u64(c, rev.get(), 0, oslots, s);
    fill_u32(c, hit.get(), 0, oslots / 32, s);
    launch(c, "q3_orders", k_q3_orders, q3_grid(c, nq), dim3(256), 0, s, (const uint32_t*)orows.as<uint32_t>(), nq,
           t.o_orderkey, otab.as<uint64_t>(), oslots - 1, dup);
    // 3. lineitems: probe and aggregate
    launch(c, "q3_lineitem", k_q3_lineitem, q3_grid(c, ceil_div(t.n_lineitem, 4)), dim3(256), 0, s, t.l_orderkey,
           t.l_shipdate, t.l_extendedprice, t.l_discount, t.n_lineitem, date, (const uint64_t*)otab.as<uint64_t>(),
           oslots - 1, rev.as<unsigned long long>(), hit.as<uint32_t>());
    // 4. groups
    Scratch gk(c, std::max<uint64_t>(nq, 1) * 4, s), gr(c, std::max<uint64_t>(nq, 1) * 8, s),
        gd(c, std::max<uint64_t>(nq, 1) * 4, s), gp(c, std::max<uint64_t>(nq, 1) * 4, s);
    launch(c, "q3_emit", k_q3_emit, dim3((uint32_t)ceil_div(oslots / 32, 256)), dim3(256), 0, s,
           (const uint64_t*)otab.as<uint64_t>(), (const int64_t*)rev.as<int64_t>(), (const uint32_t*)hit.as<uint32_t>(),
           oslots / 32, t.o_orderdate,
           t.o_shippriority, cnt + 1, gk.as<uint32_t>(), gr.as<int64_t>(), gd.as<int32_t>(), gp.as<int32_t>());
    uint64_t hv[3] = {0, 0, 0};
    read_small(c, hv, cnt, 24, s);
    PD_REQUIRE((uint32_t)hv[2] == 0, PIMDAL_E_INVALID, "o_orderkey values must be unique");
    const uint64_t g = hv[1];
    if (g == 0 || cap == 0) return g;
    // 5a. a few groups asked for: the candidates of the first `cap`, sorted in one block
    if (cap < g && cap <= Q3_TCAND && c->tune_get("q3_topk", 1)) {
        // mm[2] | thr[4] | hist[Q3_TBINS] | cand[g] | the per-block (min, max) partials
        const dim3 gg = dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(g, 256), 1024)));
        Scratch ts(c, 16 + 16 + Q3_TBINS * 4 + (uint64_t)g * 4 + 8 + (uint64_t)gg.x * 16, s);
        unsigned long long* mm = ts.as<unsigned long long>();
        uint32_t* thr = reinterpret_cast<uint32_t*>(mm + 2);   // [0] bin, [1] candidates, [2] listed
        uint32_t* hist = thr + 4;
        uint32_t* cand = hist + Q3_TBINS;
        unsigned long long* part =
            reinterpret_cast<unsigned long long*>(ts.as<char>() + ((32 + Q3_TBINS * 4 + g * 4 + 7) & ~7ull));
        fill_u32(c, thr, 0, 4 + Q3_TBINS, s);
        launch(c, "q3_topk", k_q3_tminmax, gg, dim3(256), 0, s, (const int64_t*)gr.as<int64_t>(), g, part);
        launch(c, "q3_topk", k_q3_tfold, dim3(1), dim3(256), 0, s, (const unsigned long long*)part, gg.x, mm);
        launch(c, "q3_topk", k_q3_thist, gg, dim3(256), 0, s, (const int64_t*)gr.as<int64_t>(), g,
               (const unsigned long long*)mm, hist);
        launch(c, "q3_topk", k_q3_tthresh, dim3(1), dim3(256), 0, s, (const uint32_t*)hist, cap, thr);
        uint32_t th[2] = {0, 0};
        read_small(c, th, thr, 8, s);
        if (th[1] <= Q3_TCAND) {
            launch(c, "q3_topk", k_q3_tcand, gg, dim3(256), 0, s, (const int64_t*)gr.as<int64_t>(), g,
                   (const unsigned long long*)mm, (const uint32_t*)thr, thr + 2, cand);
            launch(c, "q3_topk", k_q3_tsort, dim3(1), dim3(1024), 0, s, (const uint32_t*)cand, th[1], (uint32_t)cap,
                   (const uint32_t*)gk.as<uint32_t>(), (const int64_t*)gr.as<int64_t>(),
                   (const int32_t*)gd.as<int32_t>(), (const int32_t*)gp.as<int32_t>(), out.orderkey, out.revenue,
                   out.orderdate, out.shippriority);
            PD_HIP(hipStreamSynchronize(s));
            return g;
        }
    }
    // 5. TPC-H order: stable sorts of a permutation, least significant key first
    Scratch pa(c, g * 4, s), pb(c, g * 4, s), ka(c, g * 4, s), kb(c, g * 4, s);
    launch(c, "q3_order", k_q3_iota, q3_grid(c, g), dim3(256), 0, s, pa.as<uint32_t>(), g);
    uint32_t* perm = pa.as<uint32_t>();
    uint32_t* perm2 = pb.as<uint32_t>();
    for (uint32_t which = 0; which < 4; which++) {
        launch(c, "q3_order", k_q3_sortkey, q3_grid(c, g), dim3(256), 0, s, which, (const uint32_t*)perm, g,
               (const uint32_t*)gk.as<uint32_t>(), (const int64_t*)gr.as<int64_t>(), (const int32_t*)gd.as<int32_t>(),
               ka.as<uint32_t>());
        // bits equal in every key are detected and skipped (dates and the
        // revenue's high word vary in a few low bits only)
        sort_pairs(c, ka.as<uint32_t>(), perm, 4, g, kb.as<uint32_t>(), perm2, -1, -1, s);
        std::swap(perm, perm2);
    }
    const uint64_t k = std::min(g, cap);
    launch(c, "q3_order", k_q3_out, q3_grid(c, k), dim3(256), 0, s, (const uint32_t*)perm, k,
           (const uint32_t*)gk.as<uint32_t>(), (const int64_t*)gr.as<int64_t>(), (const int32_t*)gd.as<int32_t>(),
           (const int32_t*)gp.as<int32_t>(), out.orderkey, out.revenue, out.orderdate, out.shippriority);
    PD_HIP(hipStreamSynchronize(s));
    return g;
}

// ===================== TPC-H Q4 shape =====================
// Reference: pimdal/main/tpc_h/query4/host/host_q4.cpp:303-330 runs four DPU
// programs with the host collecting and redistributing rows between them:
// kernel_q4_1.c:161-201 selects the orders with date_start <= o_orderdate <
// date_end, kernel_q4_2.c:104-140 the lineitems with l_commitdate <
// l_receiptdate, kernel_q4_3.c:144-209 hash-joins them on orderkey and keeps
// one row per order (group by orderkey), kernel_q4_4.c:72-99 counts per
// o_orderpriority and host_q4.cpp:211-231 sums the devices' counts and
// orders them by priority.  Plan on gfx950:
//  1. k_q4_osel      the qualifying orders' rows listed (one pass);
//  2. k_q3_orders    the orders table (orderkey, order row) sized for them;
//  3. k_q4_bits / k_q4_lineitem  a filter bitmap of the qualifying orders'
//                    keys (at most 2 MB, L2-resident), then one pass over the
//                    lineitems: commitdate < receiptdate, the filter bit, the
//                    table probe, and a hit sets the slot's hit bit (the
//                    EXISTS: an order counts once however many lineitems match);
//  4. k_q4_count     hit slots counted per priority code (LDS counters, one
//                    device atomic per code per block);
//  5. k_q4_emit      the codes with a count, in code order (a sorted
//                    dictionary keeps the priority strings' order).

__global__ __launch_bounds__(256) void k_q4_osel(const int32_t* __restrict__ odate, uint64_t n, int32_t lo, int32_t hi,
                                                 unsigned long long* __restrict__ ctr, uint32_t* __restrict__ rows) {
    constexpr int PER = Q3_CHUNK / 256;   // 32 rows per thread, bit j: row r0 + j * 256 + t
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_base;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t r0 = (uint64_t)blockIdx.x * Q3_CHUNK;
    if (r0 >= n) return;   // no rows (the loads below are clamped to row n - 1)
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        // loads clamped to the last row, so all are in flight together
        const uint64_t i = r0 + (uint64_t)j * 256 + t;
        const int32_t d = odate[i < n ? i : n - 1];
        if (i < n && d >= lo && d < hi) bits |= 1u << j;
    }
    const uint32_t cnt = __popc(bits);
    const uint32_t inc = pd::wave_incl_scan(cnt);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t before = inc - cnt, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        before += (i < (int)w) ? s_w[i] : 0u;
        tot += s_w[i];
    }
    if (t == 0) s_base = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0ull;
    __syncthreads();
    uint64_t o = s_base + before;
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1;
        rows[o++] = (uint32_t)(r0 + (uint64_t)j * 256 + t);
    }
}

// Filter bit of an orderkey (the high half of the table's hash, so it is
// independent of the slot).
__device__ __forceinline__ uint32_t q4_bit(uint32_t key, uint32_t bmask) {
    return (uint32_t)(pd_mix64_dev(key) >> 32) & bmask;
}

// The qualifying orders' filter bits (a bitmap of at most 2 MB that stays
// in L2, in front of the table).
__global__ void k_q4_bits(const uint32_t* __restrict__ rows, uint64_t nq, const int32_t* __restrict__ okey,
                          uint32_t* __restrict__ bits, uint32_t bmask) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nq; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = q4_bit((uint32_t)okey[rows[j]], bmask);
        atomicOr(&bits[b >> 5], 1u << (b & 31));
    }
}

// Lineitems with l_commitdate < l_receiptdate whose order is in the table
// set the slot's hit bit.  Most lineitems' orders do not qualify; the
// filter bitmap (an L2 hit) turns those away before the table (a random
// line from MALL each).  Four rows per thread, all loads of a step in
// flight together.
__global__ __launch_bounds__(256) void k_q4_lineitem(const int32_t* __restrict__ lkey,
                                                     const int32_t* __restrict__ lcommit,
                                                     const int32_t* __restrict__ lreceipt, uint64_t n,
                                                     const uint32_t* __restrict__ bits, uint32_t bmask,
                                                     const uint64_t* __restrict__ otab, uint64_t omask,
                                                     uint32_t* __restrict__ hit) {
    constexpr int R = 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; base < n; base += stride * R) {
        uint32_t key[R], fb[R], fw[R];
        bool want[R];
        uint64_t h[R], e[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            // loads clamped to the last row, so none waits behind a branch
            const uint64_t i = base + (uint64_t)r * stride;
            const uint64_t ci = i < n ? i : n - 1;
            const int32_t dc = lcommit[ci], dr = lreceipt[ci];
            const uint32_t k = (uint32_t)lkey[ci];
            want[r] = i < n && dc < dr;
            key[r] = want[r] ? k : 0u;
            fb[r] = q4_bit(key[r], bmask);
        }
#pragma unroll
        for (int r = 0; r < R; r++) fw[r] = want[r] ? bits[fb[r] >> 5] : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) {
            want[r] = want[r] && ((fw[r] >> (fb[r] & 31)) & 1u);
            h[r] = q3_slot(key[r], omask);
        }
#pragma unroll
        for (int r = 0; r < R; r++) e[r] = want[r] ? otab[h[r]] : 0ull;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!want[r]) continue;
            while (e[r] != 0 && (uint32_t)e[r] != key[r]) {
                h[r] = (h[r] + 1) & omask;
                e[r] = otab[h[r]];
            }
            if (e[r] == 0) continue;
            const uint32_t bit = 1u << (h[r] & 31);
            if (!(hit[h[r] >> 5] & bit)) atomicOr(&hit[h[r] >> 5], bit);
        }
    }
}

// Hit slots per priority code: thread = one hit word; LDS counters, then
// one device add per code with a count in the block.
__global__ __launch_bounds__(256) void k_q4_count(const uint64_t* __restrict__ otab, const uint32_t* __restrict__ hit,
                                                  uint64_t words, const uint8_t* __restrict__ oprio,
                                                  unsigned long long* __restrict__ cnt) {
    __shared__ uint32_t s_c[256];
    const uint32_t t = threadIdx.x;
    s_c[t] = 0;
    __syncthreads();
    const uint64_t wd = (uint64_t)blockIdx.x * 256 + t;
    uint32_t bits = wd < words ? hit[wd] : 0u;
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1;
        const uint64_t e = otab[wd * 32 + (uint64_t)j];
        atomicAdd(&s_c[oprio[(e >> 32) - 1]], 1u);
    }
    __syncthreads();
    if (s_c[t]) atomicAdd(&cnt[t], (unsigned long long)s_c[t]);
}

// Codes with a count, in code order: thread t = code t; at most cap rows
// written; meta[0] = the number of such codes.
__global__ __launch_bounds__(256) void k_q4_emit(const unsigned long long* __restrict__ cnt, uint8_t* __restrict__ oprio,
                                                 int64_t* __restrict__ ocount, uint64_t cap,
                                                 unsigned long long* __restrict__ meta) {
    __shared__ uint32_t s_w[4];
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const unsigned long long c = cnt[t];
    const uint32_t f = c ? 1u : 0u;
    const uint32_t inc = pd::wave_incl_scan(f);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t before = inc - f, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        before += (i < (int)w) ? s_w[i] : 0u;
        tot += s_w[i];
    }
    if (f && before < cap) {
        oprio[before] = (uint8_t)t;
        ocount[before] = (int64_t)c;
    }
    if (t == 0) meta[0] = tot;
}

uint64_t tpch_q4(pimdal_ctx* c, const pimdal_q4_tables& t, int32_t date_start, int32_t date_end, uint8_t* out_prio,
                 int64_t* out_count, uint64_t cap, hipStream_t s) {
    PD_REQUIRE(t.o_orderkey && t.o_orderdate && t.o_orderpriority && t.l_orderkey && t.l_commitdate &&
                   t.l_receiptdate,
               PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(!cap || (out_prio && out_count), PIMDAL_E_INVALID, "missing output column");
    PD_REQUIRE(t.n_orders < (1ull << 30) && t.n_lineitem < (1ull << 40), PIMDAL_E_INVALID,
               "tables too large (at most 2^30-1 orders)");
    // small: [0] qualifying orders, [1] emitted codes, [2] repeated-key flag, [3..259) counts per code
    Scratch small(c, (3 + 256) * 8, s);
    fill_u64(c, small.get(), 0, 3 + 256, s);
    unsigned long long* meta = small.as<unsigned long long>();
    uint32_t* dup = reinterpret_cast<uint32_t*>(meta + 2);
    unsigned long long* cnt = meta + 3;
    // 1. qualifying orders, 2. their table
    Scratch orows(c, std::max<uint64_t>(t.n_orders, 1) * 4, s);
    launch(c, "q4_osel", k_q4_osel, dim3((uint32_t)std::max<uint64_t>(1, ceil_div(t.n_orders, Q3_CHUNK))), dim3(256), 0,
           s, t.o_orderdate, t.n_orders, date_start, date_end, meta, orows.as<uint32_t>());
    uint64_t nq = 0;
    read_small(c, &nq, meta, 8, s);
    const uint64_t oslots = q3_pow2_slots(nq);
    // filter bits: about 32 per qualifying order, 2^16 to 2^24 (2 MB)
    uint64_t nbits = 1ull << 16;
    while (nbits < 32 * nq && nbits < (1ull << 24)) nbits <<= 1;
    Scratch otab(c, oslots * 8, s), hit(c, oslots / 8, s), fbits(c, nbits / 8, s);
    fill_u64(c, otab.get(), 0, oslots, s);
    fill_u32(c, hit.get(), 0, oslots / 32, s);
    fill_u32(c, fbits.get(), 0, nbits / 32, s);
    launch(c, "q4_orders", k_q3_orders, q3_grid(c, nq), dim3(256), 0, s, (const uint32_t*)orows.as<uint32_t>(), nq,
           t.o_orderkey, otab.as<uint64_t>(), oslots - 1, dup);
    launch(c, "q4_bits", k_q4_bits, q3_grid(c, nq), dim3(256), 0, s, (const uint32_t*)orows.as<uint32_t>(), nq,
           t.o_orderkey, fbits.as<uint32_t>(), (uint32_t)(nbits - 1));
    // 3. lineitems mark their orders
    launch(c, "q4_lineitem", k_q4_lineitem, q3_grid(c, ceil_div(t.n_lineitem, 4)), dim3(256), 0, s, t.l_orderkey,
           t.l_commitdate, t.l_receiptdate, t.n_lineitem, (const uint32_t*)fbits.as<uint32_t>(),
           (uint32_t)(nbits - 1), (const uint64_t*)otab.as<uint64_t>(), oslots - 1, hit.as<uint32_t>());
    // 4. counts per priority, 5. in priority order
    launch(c, "q4_count", k_q4_count, dim3((uint32_t)ceil_div(oslots / 32, 256)), dim3(256), 0, s,
           (const uint64_t*)otab.as<uint64_t>(), (const uint32_t*)hit.as<uint32_t>(), oslots / 32, t.o_orderpriority,
           cnt);
    launch(c, "q4_emit", k_q4_emit, dim3(1), dim3(256), 0, s, (const unsigned long long*)cnt, out_prio, out_count, cap,
           meta + 1);
    uint64_t hv[3] = {0, 0, 0};
    read_small(c, hv, meta, 24, s);
    PD_REQUIRE((uint32_t)hv[2] == 0, PIMDAL_E_INVALID, "o_orderkey values must be unique");
    return hv[1];
}

// ===================== TPC-H Q5 shape =====================
// Reference: pimdal/main/tpc_h/query5/host/host_q5.cpp:566-640 runs seven
// DPU programs with three host redistributions: kernel_q5_1.c selects the
// region by name, its nations and their customers (custkey, nationkey);
// kernel_q5_2.c the orders of [date_start, date_end); kernel_q5_3.c joins
// them on custkey (orderkey, nationkey); kernel_q5_4.c / kernel_q5_5.c join
// the lineitems on orderkey; kernel_q5_6.c / kernel_q5_7.c join the
// suppliers on (suppkey, nationkey) and sum l_extendedprice * (100 -
// l_discount) per nation; host_q5.cpp:459-500 sums the devices' rows and
// orders by revenue.  Plan on gfx950:
//  1. k_q5_nations  the region's nations as flags (nation keys < 1024);
//  2. k_q5_cust     the region's customers in a table custkey -> nation;
//  3. k_q5_osel / k_q5_orders   orders of the date range whose customer is
//                   in the table, listed, then tabled orderkey -> nation,
//                   with a filter bitmap of their keys (as Q4's);
//  4. k_q5_supp     suppliers tabled suppkey -> nation (all of them);
//  5. k_q5_lineitem one pass: filter bit, orders table, suppliers table,
//                   and when the two nations agree the revenue and a row
//                   count added per nation in LDS, one device add per
//                   nation per block;
//  6. k_q5_emit     nations with a row, by revenue descending then key.
constexpr uint32_t Q5_NATIONS = 1024;

// flags[k] = 1 for the nations of a region named `name`; flags[Q5_NATIONS]
// = 1 when a nation key is out of range.  One block.
__global__ __launch_bounds__(256) void k_q5_nations(const int32_t* __restrict__ rkey, const uint8_t* __restrict__ rname,
                                                    uint64_t nr, const int32_t* __restrict__ nkey,
                                                    const int32_t* __restrict__ nreg, uint64_t nn, uint32_t name,
                                                    uint32_t* __restrict__ flags) {
    for (uint64_t i = threadIdx.x; i < nn; i += blockDim.x) {
        const int32_t k = nkey[i];
        if (k < 0 || k >= (int32_t)Q5_NATIONS) {
            flags[Q5_NATIONS] = 1;
            continue;
        }
        bool in = false;
        for (uint64_t r = 0; r < nr && !in; r++) in = rname[r] == name && rkey[r] == nreg[i];
        if (in) flags[k] = 1;
    }
}

// Table slot word: (nation + 1) << 32 | key (0 = empty).  Repeated keys
// set *dup.
__device__ __forceinline__ void q5_insert(uint64_t* tab, uint64_t mask, uint32_t key, uint32_t nation, uint32_t* dup) {
    const uint64_t want = ((uint64_t)(nation + 1) << 32) | key;
    uint64_t h = q3_slot(key, mask);
    while (true) {
        const uint64_t prev = atomicCAS((unsigned long long*)&tab[h], 0ull, (unsigned long long)want);
        if (prev == 0) break;
        if ((uint32_t)prev == key) {
            atomicOr(dup, 1u);
            break;
        }
        h = (h + 1) & mask;
    }
}

// The nation of `key` in a table, or -1.
__device__ __forceinline__ int32_t q5_find(const uint64_t* tab, uint64_t mask, uint32_t key) {
    uint64_t h = q3_slot(key, mask);
    while (true) {
        const uint64_t e = tab[h];
        if (e == 0) return -1;
        if ((uint32_t)e == key) return (int32_t)(e >> 32) - 1;
        h = (h + 1) & mask;
    }
}

// The region's customers: keys below `dn` in a direct array (entry = nation
// + 1, 0 = none; dense TPC-H custkeys make it a few MB that the caches
// hold, where the hash table is a random MALL read per probe), larger keys
// in the hash table.  A key met twice sets *dup either way.
__global__ void k_q5_cust(const int32_t* __restrict__ ckey, const int32_t* __restrict__ cnat, uint64_t n,
                          const uint32_t* __restrict__ flags, uint32_t* __restrict__ cdir, uint32_t dn,
                          uint64_t* __restrict__ tab, uint64_t mask, uint32_t* __restrict__ dup) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t nat = cnat[i];
        if (nat < 0 || nat >= (int32_t)Q5_NATIONS || !flags[nat]) continue;
        const uint32_t k = (uint32_t)ckey[i];
        if (k < dn) {
            if (atomicCAS(&cdir[k], 0u, (uint32_t)nat + 1) != 0) atomicOr(dup, 1u);
            continue;
        }
        q5_insert(tab, mask, k, (uint32_t)nat, dup);
    }
}

// The nation of customer `key`, or -1.
__device__ __forceinline__ int32_t q5_cust_nation(const uint32_t* cdir, uint32_t dn, const uint64_t* tab,
                                                  uint64_t mask, uint32_t key) {
    return key < dn ? (int32_t)cdir[key] - 1 : q5_find(tab, mask, key);
}

// Orders of [lo, hi) whose customer is in the table: their rows listed
// (one counter atomic per 8192-row block, as k_q3_osel).
__global__ __launch_bounds__(256) void k_q5_osel(const int32_t* __restrict__ odate, const int32_t* __restrict__ ocust,
                                                 uint64_t n, int32_t lo, int32_t hi, const uint32_t* __restrict__ cdir,
                                                 uint32_t dn, const uint64_t* __restrict__ ctab, uint64_t cmask,
                                                 unsigned long long* __restrict__ ctr, uint32_t* __restrict__ rows) {
    constexpr int PER = Q3_CHUNK / 256;
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_base;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    const uint64_t r0 = (uint64_t)blockIdx.x * Q3_CHUNK;
    if (r0 >= n) return;   // no rows (the loads below are clamped to row n - 1)
    uint32_t bits = 0;
#pragma unroll
    for (int j0 = 0; j0 < PER; j0 += 4) {
        bool cand[4];
        uint32_t key[4];
        uint64_t h[4], e[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // loads clamped to the last row, so none waits behind a branch
            const uint64_t i = r0 + (uint64_t)(j0 + q) * 256 + t;
            const uint64_t ci = i < n ? i : n - 1;
            const int32_t d = odate[ci];
            const uint32_t cu = (uint32_t)ocust[ci];
            cand[q] = i < n && d >= lo && d < hi;
            key[q] = cu;
            h[q] = q3_slot(cu, cmask);
        }
        uint32_t dv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            dv[q] = cand[q] && key[q] < dn ? cdir[key[q]] : 0u;
            e[q] = cand[q] && key[q] >= dn ? ctab[h[q]] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!cand[q]) continue;
            if (key[q] < dn) {
                if (dv[q] != 0) bits |= 1u << (j0 + q);
                continue;
            }
            while (e[q] != 0 && (uint32_t)e[q] != key[q]) {
                h[q] = (h[q] + 1) & cmask;
                e[q] = ctab[h[q]];
            }
            if (e[q] != 0) bits |= 1u << (j0 + q);
        }
    }
    const uint32_t cnt = __popc(bits);
    const uint32_t inc = pd::wave_incl_scan(cnt);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t before = inc - cnt, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        before += (i < (int)w) ? s_w[i] : 0u;
        tot += s_w[i];
    }
    if (t == 0) s_base = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0ull;
    __syncthreads();
    uint64_t o = s_base + before;
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1;
        rows[o++] = (uint32_t)(r0 + (uint64_t)j * 256 + t);
    }
}

// The listed orders tabled orderkey -> their customer's nation, with their
// filter bits.
__global__ void k_q5_orders(const uint32_t* __restrict__ rows, uint64_t nq, const int32_t* __restrict__ okey,
                            const int32_t* __restrict__ ocust, const uint32_t* __restrict__ cdir, uint32_t dn,
                            const uint64_t* __restrict__ ctab, uint64_t cmask, uint64_t* __restrict__ otab,
                            uint64_t omask, uint32_t* __restrict__ bits, uint32_t bmask, uint32_t* __restrict__ dup) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nq; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = rows[j];
        const uint32_t k = (uint32_t)okey[i];
        const int32_t nat = q5_cust_nation(cdir, dn, ctab, cmask, (uint32_t)ocust[i]);
        q5_insert(otab, omask, k, (uint32_t)nat, dup);
        const uint32_t b = q4_bit(k, bmask);
        atomicOr(&bits[b >> 5], 1u << (b & 31));
    }
}

__global__ void k_q5_supp(const int32_t* __restrict__ skey, const int32_t* __restrict__ snat, uint64_t n,
                          uint64_t* __restrict__ tab, uint64_t mask, uint32_t* __restrict__ dup) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t nat = snat[i];
        // a supplier outside the nation range matches no order's nation
        q5_insert(tab, mask, (uint32_t)skey[i], nat >= 0 && nat < (int32_t)Q5_NATIONS ? (uint32_t)nat : Q5_NATIONS, dup);
    }
}

// Lineitems: filter bit, the order's nation, the supplier's nation; equal
// nations add price * (100 - discount) and one row to the nation's LDS
// accumulators.  Four rows per thread, the loads of a step in flight.
__global__ __launch_bounds__(256) void k_q5_lineitem(const int32_t* __restrict__ lkey,
                                                     const int32_t* __restrict__ lsupp,
                                                     const int64_t* __restrict__ lprice,
                                                     const int64_t* __restrict__ ldisc, uint64_t n,
                                                     const uint32_t* __restrict__ bits, uint32_t bmask,
                                                     const uint64_t* __restrict__ otab, uint64_t omask,
                                                     const uint64_t* __restrict__ stab, uint64_t smask,
                                                     unsigned long long* __restrict__ rev,
                                                     unsigned long long* __restrict__ rows) {
    constexpr int R = 4;
    __shared__ unsigned long long s_rev[Q5_NATIONS];
    __shared__ uint32_t s_rows[Q5_NATIONS];
    for (uint32_t i = threadIdx.x; i < Q5_NATIONS; i += 256) {
        s_rev[i] = 0;
        s_rows[i] = 0;
    }
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; base < n; base += stride * R) {
        uint32_t key[R], fb[R], fw[R];
        bool want[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t i = base + (uint64_t)r * stride;
            want[r] = i < n;
            key[r] = (uint32_t)lkey[want[r] ? i : n - 1];   // clamped: no load waits behind a branch
            fb[r] = q4_bit(key[r], bmask);
        }
#pragma unroll
        for (int r = 0; r < R; r++) fw[r] = want[r] ? bits[fb[r] >> 5] : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!want[r] || !((fw[r] >> (fb[r] & 31)) & 1u)) continue;
            const int32_t on = q5_find(otab, omask, key[r]);
            if (on < 0) continue;
            const uint64_t i = base + (uint64_t)r * stride;
            const int32_t sn = q5_find(stab, smask, (uint32_t)lsupp[i]);
            if (sn != on) continue;
            atomicAdd(&s_rev[on], (unsigned long long)(lprice[i] * (100 - ldisc[i])));
            atomicAdd(&s_rows[on], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < Q5_NATIONS; i += 256) {
        if (s_rows[i]) {
            atomicAdd(&rev[i], s_rev[i]);
            atomicAdd(&rows[i], (unsigned long long)s_rows[i]);
        }
    }
}

// Nations with a row, revenue descending then nationkey ascending: the
// nations with a row are listed (any order), then each one's rank is the
// number of listed nations before it.  One block of 1024.
__global__ __launch_bounds__(1024) void k_q5_emit(const unsigned long long* __restrict__ rev,
                                                  const unsigned long long* __restrict__ rows, int32_t* __restrict__ onat,
                                                  int64_t* __restrict__ orev, uint64_t cap,
                                                  unsigned long long* __restrict__ meta) {
    __shared__ int64_t s_rev[Q5_NATIONS];
    __shared__ uint32_t s_key[Q5_NATIONS];
    __shared__ uint32_t s_n;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_n = 0;
    __syncthreads();
    if (rows[t]) {
        const uint32_t j = atomicAdd(&s_n, 1u);
        s_rev[j] = (int64_t)rev[t];
        s_key[j] = t;
    }
    __syncthreads();
    const uint32_t n = s_n;
    if (t < n) {
        const int64_t r = s_rev[t];
        const uint32_t k = s_key[t];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < n; j++) rank += (s_rev[j] > r || (s_rev[j] == r && s_key[j] < k)) ? 1u : 0u;
        if (rank < cap) {
            onat[rank] = (int32_t)k;
            orev[rank] = r;
        }
    }
    if (t == 0) meta[0] = n;
}

uint64_t tpch_q5(pimdal_ctx* c, const pimdal_q5_tables& t, uint32_t region_name, int32_t date_start, int32_t date_end,
                 int32_t* out_nation, int64_t* out_rev, uint64_t cap, hipStream_t s) {
    PD_REQUIRE(t.r_regionkey && t.r_name && t.n_nationkey && t.n_regionkey && t.c_custkey && t.c_nationkey &&
                   t.o_orderkey && t.o_custkey && t.o_orderdate && t.l_orderkey && t.l_suppkey && t.l_extendedprice &&
                   t.l_discount && t.s_suppkey && t.s_nationkey,
               PIMDAL_E_INVALID, "missing input column");
    PD_REQUIRE(!cap || (out_nation && out_rev), PIMDAL_E_INVALID, "missing output column");
    PD_REQUIRE(t.n_orders < (1ull << 30) && t.n_customer < (1ull << 30) && t.n_supplier < (1ull << 30) &&
                   t.n_lineitem < (1ull << 40) && t.n_region <= 4096 && t.n_nation <= 4096,
               PIMDAL_E_INVALID, "tables too large");
    // small: flags[Q5_NATIONS + 1] u32 | dup u32 | pad | counters: [0] listed orders, [1] nations out,
    // rev[Q5_NATIONS], rows[Q5_NATIONS]
    const uint64_t fw = Q5_NATIONS + 2;
    Scratch small(c, fw * 4 + (2 + 2 * (uint64_t)Q5_NATIONS) * 8 + 8, s);
    fill_u64(c, small.get(), 0, (fw * 4 + (2 + 2 * (uint64_t)Q5_NATIONS) * 8 + 8) / 8, s);
    uint32_t* flags = small.as<uint32_t>();
    uint32_t* dup = flags + Q5_NATIONS + 1;
    unsigned long long* meta = reinterpret_cast<unsigned long long*>(small.as<char>() + ((fw * 4 + 7) & ~7ull));
    unsigned long long* rev = meta + 2;
    unsigned long long* rows = rev + Q5_NATIONS;
    // 1. nations, 2. customers
    launch(c, "q5_nations", k_q5_nations, dim3(1), dim3(256), 0, s, t.r_regionkey, t.r_name, t.n_region, t.n_nationkey,
           t.n_regionkey, t.n_nation, region_name, flags);
    const uint64_t cslots = q3_pow2_slots(t.n_customer);
    const uint32_t dn = q3_direct_keys(t.n_customer, 22);
    Scratch ctab(c, cslots * 8, s), cdir(c, (uint64_t)dn * 4, s);
    fill_u64(c, ctab.get(), 0, cslots, s);
    fill_u32(c, cdir.get(), 0, dn, s);
    launch(c, "q5_cust", k_q5_cust, q3_grid(c, t.n_customer), dim3(256), 0, s, t.c_custkey, t.c_nationkey, t.n_customer,
           (const uint32_t*)flags, cdir.as<uint32_t>(), dn, ctab.as<uint64_t>(), cslots - 1, dup);
    // 3. orders of the range with a customer of the region
    Scratch orows(c, std::max<uint64_t>(t.n_orders, 1) * 4, s);
    launch(c, "q5_osel", k_q5_osel, dim3((uint32_t)std::max<uint64_t>(1, ceil_div(t.n_orders, Q3_CHUNK))), dim3(256), 0,
           s, t.o_orderdate, t.o_custkey, t.n_orders, date_start, date_end, (const uint32_t*)cdir.as<uint32_t>(), dn,
           (const uint64_t*)ctab.as<uint64_t>(), cslots - 1, meta, orows.as<uint32_t>());
    uint32_t hflags[2] = {0, 0};
    read_small(c, hflags, flags + Q5_NATIONS, 8, s);
    PD_REQUIRE(hflags[0] == 0, PIMDAL_E_UNSUPPORTED, "nation keys must lie in [0, 1024)");
    PD_REQUIRE(hflags[1] == 0, PIMDAL_E_INVALID, "c_custkey values of the region's customers must be unique");
    uint64_t nq = 0;
    read_small(c, &nq, meta, 8, s);
    const uint64_t oslots = q3_pow2_slots(nq);
    uint64_t nbits = 1ull << 16;
    while (nbits < 32 * nq && nbits < (1ull << 24)) nbits <<= 1;
    const uint64_t sslots = q3_pow2_slots(t.n_supplier);
    Scratch otab(c, oslots * 8, s), fbits(c, nbits / 8, s), stab(c, sslots * 8, s);
    fill_u64(c, otab.get(), 0, oslots, s);
    fill_u32(c, fbits.get(), 0, nbits / 32, s);
    fill_u64(c, stab.get(), 0, sslots, s);
    launch(c, "q5_orders", k_q5_orders, q3_grid(c, nq), dim3(256), 0, s, (const uint32_t*)orows.as<uint32_t>(), nq,
           t.o_orderkey, t.o_custkey, (const uint32_t*)cdir.as<uint32_t>(), dn, (const uint64_t*)ctab.as<uint64_t>(),
           cslots - 1, otab.as<uint64_t>(), oslots - 1, fbits.as<uint32_t>(), (uint32_t)(nbits - 1), dup);
    // 4. suppliers, 5. lineitems
    launch(c, "q5_supp", k_q5_supp, q3_grid(c, t.n_supplier), dim3(256), 0, s, t.s_suppkey, t.s_nationkey, t.n_supplier,
           stab.as<uint64_t>(), sslots - 1, dup);
    launch(c, "q5_lineitem", k_q5_lineitem, q3_grid(c, ceil_div(t.n_lineitem, 4)), dim3(256), 0, s, t.l_orderkey,
           t.l_suppkey, t.l_extendedprice, t.l_discount, t.n_lineitem, (const uint32_t*)fbits.as<uint32_t>(),
           (uint32_t)(nbits - 1), (const uint64_t*)otab.as<uint64_t>(), oslots - 1,
           (const uint64_t*)stab.as<uint64_t>(), sslots - 1, rev, rows);
    // 6. nations by revenue
    launch(c, "q5_emit", k_q5_emit, dim3(1), dim3(Q5_NATIONS), 0, s, (const unsigned long long*)rev,
           (const unsigned long long*)rows, out_nation, out_rev, cap, meta + 1);
    uint64_t g = 0;
    read_small(c, &g, meta + 1, 8, s);
    uint32_t d2 = 0;
    read_small(c, &d2, dup, 4, s);
    PD_REQUIRE(d2 == 0, PIMDAL_E_INVALID, "orderkey and suppkey values must be unique");
    return g;
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_tpch_q3_dev(pimdal_ctx* c, const pimdal_q3_tables* tables, uint32_t segment, int32_t date,
                                  const pimdal_q3_result* out, uint64_t capacity, uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && tables && out && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = tpch_q3(c, *tables, segment, date, *out, capacity, c->stream);
    PD_API_END
}

extern "C" int pimdal_tpch_q4_dev(pimdal_ctx* c, const pimdal_q4_tables* tables, int32_t date_start, int32_t date_end,
                                  uint8_t* out_priority, int64_t* out_count, uint64_t capacity, uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && tables && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = tpch_q4(c, *tables, date_start, date_end, out_priority, out_count, capacity, c->stream);
    PD_API_END
}

extern "C" int pimdal_tpch_q5_dev(pimdal_ctx* c, const pimdal_q5_tables* tables, uint32_t region_name,
                                  int32_t date_start, int32_t date_end, int32_t* out_nationkey, int64_t* out_revenue,
                                  uint64_t capacity, uint64_t* n_groups) {
    PD_API_BEGIN
    PD_REQUIRE(c && tables && n_groups, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    *n_groups = tpch_q5(c, *tables, region_name, date_start, date_end, out_nationkey, out_revenue, capacity, c->stream);
    PD_API_END
}
