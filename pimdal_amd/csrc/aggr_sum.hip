// aggr_sum.hip — SUM and COUNT per uint32 key: the hash and sort
// aggregations' device paths.
//
// Reference: reference/db_operators/aggregate.cpp:40-151 (hash_aggr:
// thread-local hash maps merged under a critical section);
// pimdal/aggregate/aggregate_hash.c:43-193 (group_kernel, hash variant: one
// shared 4096-entry table, spill rounds for the rows it cannot place) and
// pimdal/aggregate/aggregate.c:82-158 (group_kernel, sort variant: a
// run-length reduce over sorted 128-element blocks, boundary runs merged
// between tasklets).
//
// Few groups.  One block per CU reads its contiguous share of the rows, 8
// rows per thread per round with the round's loads in flight together, into
// an 8192-slot LDS table (key, int64 sum, uint32 count; 4-slot buckets, two
// candidate buckets per key, keys inserted by compare-and-swap; a round
// first reads both buckets of every row, 16 bytes each, and only rows found
// in neither walk on), then adds its table into a 65536-slot HBM table.  A block that meets more than AS_LIMIT distinct keys, or an
// HBM table past AS_GLIMIT keys, raises a flag that stops every block; the
// caller then takes its many-groups path.  Groups leave the HBM table in
// slot order (hash aggregation) or in key order (sort aggregation: a radix
// sort of the group keys and a gather).  Key 0xFFFFFFFF, the table's empty
// marker, is summed in a slot of its own.
//
// Sorted rows (sort aggregation with many groups).  The (key, value) pairs
// sorted by the radix sort, then one pass over 8192-row tiles: run heads
// counted per tile, group numbers from a decoupled look-back over the head
// counts, a segmented scan of (sum, count) across the tile, and every run
// that starts in the tile stored by the thread that holds its last row
// there.  The piece of a run carried over from an earlier tile is kept per
// tile and added by a second kernel, after the first has stored every run.
#include <algorithm>
#include <vector>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr uint32_t AS_BLOCK = 1024;
constexpr uint32_t AS_U = 8;                 // rows per thread per round
constexpr uint32_t AS_TS = 8192;             // LDS table slots
constexpr uint32_t AS_LIMIT = AS_TS / 2;     // distinct keys a block may meet
constexpr uint32_t AS_GT = 65536;            // HBM table slots
constexpr uint32_t AS_GLIMIT = AS_GT / 2;    // distinct keys in all
constexpr uint32_t AS_R = 8;                 // HBM table copies
constexpr uint32_t AS_EMPTY = 0xFFFFFFFFu;

// HBM tables: AS_R copies (block b adds into copy b % AS_R, so each key's
// words take a few hundred device-scope atomic adds spread over 8 addresses
// instead of on one), folded into copy 0 before the groups are emitted.
// Copy r: gkey + r * AS_GT, gsum + r * (AS_GT + 1) (word AS_GT: the sum of
// key 0xFFFFFFFF), gcnt + r * AS_GT.  meta words: [0] gave up, [2] rows with
// key 0xFFFFFFFF, [3] emit counter, [4 + r] keys in copy r.

__device__ __forceinline__ uint32_t lds_ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_flag(uint32_t* p) {
    __hip_atomic_store(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS table: 4-slot buckets filled in slot order.  A key lives in its first
// bucket b1 while that has room, else in its second bucket b2, else in the
// first bucket after b2 with room (buckets only fill, so a key once placed
// stays found).  as_slot_slow returns key k's slot, inserting it when
// absent, or AS_TS + 1 once the block has given up.  Each pass either finds
// a bucket full for good (and moves on) or a slot of it filled for good, so
// the loop ends within 5 * (AS_NB + 2) passes.
constexpr uint32_t AS_NB = AS_TS / 4;

__device__ __forceinline__ uint32_t as_b1(uint32_t k) { return slot_hash(k) & (AS_NB - 1); }
__device__ __forceinline__ uint32_t as_b2(uint32_t k) {
    const uint32_t b1 = as_b1(k);
    const uint32_t d = (uint32_t)(((uint64_t)k * 0xD6E8FEB86659FD93ull) >> 53) | 1u;   // odd step: b2 != b1
    return (b1 + d) & (AS_NB - 1);
}

__device__ uint32_t as_slot_slow(uint32_t* s_key, uint32_t* s_used, uint32_t* s_over, uint32_t k) {
    const uint32_t b1 = as_b1(k), b2 = as_b2(k);
    uint32_t b = b1, stage = 0;   // stage 0: b1, 1: b2, 2: the buckets after b2
    for (uint32_t pass = 0; pass < 5 * (AS_NB + 2); pass++) {
        if (lds_ld(s_over)) return AS_TS + 1;
        uint32_t kk[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) kk[j] = lds_ld(&s_key[b * 4 + j]);
        uint32_t e = 4;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            if (kk[j] == k) return b * 4 + j;
            if (kk[j] == AS_EMPTY && e == 4) e = j;
        }
        if (e == 4) {   // full: the next place k may be
            b = stage == 0 ? b2 : (b + 1) & (AS_NB - 1);
            stage = stage == 0 ? 1 : 2;
            continue;
        }
        const uint32_t prev = atomicCAS(&s_key[b * 4 + e], AS_EMPTY, k);
        if (prev == AS_EMPTY) {
            if (atomicAdd(s_used, 1u) >= AS_LIMIT) lds_flag(s_over);
            return b * 4 + e;
        }
        if (prev == k) return b * 4 + e;
    }
    lds_flag(s_over);
    return AS_TS + 1;
}

// Adds (sum, cnt) to key kk's entry of one HBM table copy, inserting it when
// absent.  False when the copy is past AS_GLIMIT keys or full.
__device__ bool as_ginsert(uint32_t* gkey, uint64_t* gsum, uint64_t* gcnt, uint64_t* nkeys, uint32_t kk, uint64_t sum,
                           uint64_t cnt) {
    uint32_t h = slot_hash(kk) & (AS_GT - 1);
    for (uint32_t probes = 0; probes < AS_GT; probes++) {
        const uint32_t cur = ld_relaxed(&gkey[h]);
        bool here = cur == kk;
        if (!here && cur == AS_EMPTY) {
            const uint32_t prev = atomicCAS(&gkey[h], AS_EMPTY, kk);
            if (prev == AS_EMPTY) {
                if (atomicAdd((unsigned long long*)nkeys, 1ull) >= AS_GLIMIT) return false;
                here = true;
            } else {
                here = prev == kk;
            }
        }
        if (here) {
            atomicAdd((unsigned long long*)&gsum[h], (unsigned long long)sum);
            atomicAdd((unsigned long long*)&gcnt[h], (unsigned long long)cnt);
            return true;
        }
        h = (h + 1) & (AS_GT - 1);
    }
    return false;
}

__global__ __launch_bounds__(AS_BLOCK) void k_as_table(const uint32_t* __restrict__ key,
                                                       const int64_t* __restrict__ val, uint64_t n, uint64_t chunk,
                                                       uint32_t* __restrict__ gkey, uint64_t* __restrict__ gsum,
                                                       uint64_t* __restrict__ gcnt, uint64_t* __restrict__ meta) {
    __shared__ __align__(16) uint32_t s_key[AS_TS];
    __shared__ uint64_t s_sum[AS_TS + 1];   // slot AS_TS: key 0xFFFFFFFF
    __shared__ uint32_t s_cnt[AS_TS + 1];
    __shared__ uint32_t s_used, s_over;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < AS_TS + 1; i += AS_BLOCK) {
        if (i < AS_TS) s_key[i] = AS_EMPTY;
        s_sum[i] = 0;
        s_cnt[i] = 0;
    }
    if (t == 0) {
        s_used = 0;
        s_over = 0;
    }
    __syncthreads();
    const uint64_t start = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = min(n, start + chunk);
    uint32_t round = 0;
    for (uint64_t i0 = start; i0 < end; i0 += (uint64_t)AS_BLOCK * AS_U, round++) {
        if (lds_ld(&s_over)) break;
        if ((round & 15) == 15 && ld_relaxed(&meta[0])) break;
        uint32_t k[AS_U];
        int64_t v[AS_U];
#pragma unroll
        for (uint32_t u = 0; u < AS_U; u++) {
            const uint64_t i = i0 + (uint64_t)u * AS_BLOCK + t;
            k[u] = i < end ? key[i] : AS_EMPTY;
            v[u] = i < end ? val[i] : 0;
        }
        // every row's two buckets read in one go; only rows found in
        // neither (new keys, or keys past b2) take the slow path
        uint32_t h[AS_U];
        uint4 q1[AS_U], q2[AS_U];
#pragma unroll
        for (uint32_t u = 0; u < AS_U; u++) {
            q1[u] = *reinterpret_cast<const uint4*>(&s_key[as_b1(k[u]) * 4]);
            q2[u] = *reinterpret_cast<const uint4*>(&s_key[as_b2(k[u]) * 4]);
        }
        bool gave_up = false;
#pragma unroll
        for (uint32_t u = 0; u < AS_U; u++) {
            const bool live = i0 + (uint64_t)u * AS_BLOCK + t < end;
            if (!live) {
                h[u] = AS_TS + 1;
                continue;
            }
            if (k[u] == AS_EMPTY) {
                h[u] = AS_TS;
                continue;
            }
            const uint32_t b1 = as_b1(k[u]) * 4, b2 = as_b2(k[u]) * 4;
            if (q1[u].x == k[u]) h[u] = b1;
            else if (q1[u].y == k[u]) h[u] = b1 + 1;
            else if (q1[u].z == k[u]) h[u] = b1 + 2;
            else if (q1[u].w == k[u]) h[u] = b1 + 3;
            else if (q2[u].x == k[u]) h[u] = b2;
            else if (q2[u].y == k[u]) h[u] = b2 + 1;
            else if (q2[u].z == k[u]) h[u] = b2 + 2;
            else if (q2[u].w == k[u]) h[u] = b2 + 3;
            else h[u] = as_slot_slow(s_key, &s_used, &s_over, k[u]);
            gave_up |= h[u] > AS_TS;
        }
        if (gave_up) break;
#pragma unroll
        for (uint32_t u = 0; u < AS_U; u++) {
            if (h[u] > AS_TS) continue;
            atomicAdd((unsigned long long*)&s_sum[h[u]], (unsigned long long)v[u]);
            atomicAdd(&s_cnt[h[u]], 1u);
        }
    }
    __syncthreads();
    if (s_over) {
        if (t == 0) st_relaxed(&meta[0], (uint64_t)1);
        return;
    }
    if (ld_relaxed(&meta[0])) return;   // another block gave up: this result is not used
    const uint32_t r = blockIdx.x % AS_R;
    uint32_t* rkey = gkey + (uint64_t)r * AS_GT;
    uint64_t* rsum = gsum + (uint64_t)r * (AS_GT + 1);
    uint64_t* rcnt = gcnt + (uint64_t)r * AS_GT;
    for (uint32_t i = t; i < AS_TS; i += AS_BLOCK) {
        const uint32_t kk = s_key[i];
        if (kk == AS_EMPTY) continue;
        if (!as_ginsert(rkey, rsum, rcnt, &meta[4 + r], kk, s_sum[i], s_cnt[i])) st_relaxed(&meta[0], (uint64_t)1);
    }
    if (t == 0 && s_cnt[AS_TS]) {
        atomicAdd((unsigned long long*)&rsum[AS_GT], (unsigned long long)s_sum[AS_TS]);
        atomicAdd((unsigned long long*)&meta[2], (unsigned long long)s_cnt[AS_TS]);
    }
}

// Folds table copies 1 .. AS_R - 1 into copy 0.
__global__ __launch_bounds__(256) void k_as_fold(uint32_t* __restrict__ gkey, uint64_t* __restrict__ gsum,
                                                 uint64_t* __restrict__ gcnt, uint64_t* __restrict__ meta) {
    const uint64_t total = (uint64_t)(AS_R - 1) * AS_GT;
    for (uint64_t x = blockIdx.x * 256ull + threadIdx.x; x < total; x += (uint64_t)gridDim.x * 256) {
        const uint32_t r = 1 + (uint32_t)(x / AS_GT), sl = (uint32_t)(x % AS_GT);
        const uint32_t kk = gkey[(uint64_t)r * AS_GT + sl];
        if (kk == AS_EMPTY) continue;
        if (!as_ginsert(gkey, gsum, gcnt, &meta[4], kk, gsum[(uint64_t)r * (AS_GT + 1) + sl],
                        gcnt[(uint64_t)r * AS_GT + sl]))
            st_relaxed(&meta[0], (uint64_t)1);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint64_t sp = 0;
        for (uint32_t r = 1; r < AS_R; r++) sp += gsum[(uint64_t)r * (AS_GT + 1) + AS_GT];
        atomicAdd((unsigned long long*)&gsum[AS_GT], (unsigned long long)sp);
    }
}

// Groups of the HBM table into (okey, osum, ocnt) in slot order, the
// 0xFFFFFFFF group (if any rows had that key) after them; with oidx, each
// group's own position too (the sort's payload).  One counter add per wave.
__global__ __launch_bounds__(256) void k_as_emit(const uint32_t* __restrict__ gkey, const uint64_t* __restrict__ gsum,
                                                 const uint64_t* __restrict__ gcnt, uint64_t* __restrict__ meta,
                                                 uint32_t* __restrict__ okey, int64_t* __restrict__ osum,
                                                 int64_t* __restrict__ ocnt, uint32_t* __restrict__ oidx, uint64_t cap) {
    for (uint32_t s0 = blockIdx.x * 256u; s0 < AS_GT; s0 += gridDim.x * 256u) {
        const uint32_t s = s0 + threadIdx.x;
        const uint32_t kk = gkey[s];
        const bool occ = kk != AS_EMPTY;
        const uint64_t m = ballot(occ);
        if (!m) continue;
        unsigned long long b = 0;
        if (lane() == 0) b = atomicAdd((unsigned long long*)&meta[3], (unsigned long long)popc64(m));
        const uint64_t o = (uint64_t)__shfl(b, 0, PD_WAVE) + popc64(m & lanemask_lt());
        if (occ && o < cap) {
            okey[o] = kk;
            osum[o] = (int64_t)gsum[s];
            ocnt[o] = (int64_t)gcnt[s];
            if (oidx) oidx[o] = (uint32_t)o;
        }
    }
}

__global__ void k_as_emit_last(const uint64_t* __restrict__ gsum, const uint64_t* __restrict__ meta,
                               uint32_t* __restrict__ okey, int64_t* __restrict__ osum, int64_t* __restrict__ ocnt,
                               uint32_t* __restrict__ oidx, uint64_t at) {
    okey[at] = AS_EMPTY;
    osum[at] = (int64_t)gsum[AS_GT];
    ocnt[at] = (int64_t)meta[2];
    if (oidx) oidx[at] = (uint32_t)at;
}

__global__ __launch_bounds__(256) void k_as_gather(uint64_t g, const uint32_t* __restrict__ idx,
                                                   const int64_t* __restrict__ s_in, const int64_t* __restrict__ c_in,
                                                   int64_t* __restrict__ osum, int64_t* __restrict__ ocnt) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < g; i += (uint64_t)gridDim.x * 256) {
        const uint32_t j = idx[i];
        osum[i] = s_in[j];
        ocnt[i] = c_in[j];
    }
}

bool aggr_sum_small(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey, int64_t* osum,
                    int64_t* ocnt, uint64_t cap, bool ordered, uint64_t* groups, hipStream_t s) {
    // gsum[AS_R][AS_GT + 1] | gcnt[AS_R][AS_GT] | meta[4 + AS_R] | gkey[AS_R][AS_GT]
    const uint64_t words = (uint64_t)AS_R * (AS_GT + 1) + (uint64_t)AS_R * AS_GT + 4 + AS_R;
    Scratch tab(c, words * 8 + (uint64_t)AS_R * AS_GT * 4, s);
    uint64_t* gsum = tab.as<uint64_t>();
    uint64_t* gcnt = gsum + (uint64_t)AS_R * (AS_GT + 1);
    uint64_t* meta = gcnt + (uint64_t)AS_R * AS_GT;
    uint32_t* gkey = reinterpret_cast<uint32_t*>(meta + 4 + AS_R);
    fill_u64(c, gsum, 0, words, s);
    fill_u32(c, gkey, AS_EMPTY, (uint64_t)AS_R * AS_GT, s);
    const uint32_t G = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(ceil_div(n, (uint64_t)AS_BLOCK * AS_U), (uint64_t)c->cu_count));
    const uint64_t chunk = ceil_div(n, G);
    launch(c, "as_table", k_as_table, dim3(G), dim3(AS_BLOCK), 0, s, key, val, n, chunk, gkey, gsum, gcnt, meta);
    if (G > 1)
        launch(c, "as_fold", k_as_fold, dim3(std::min<uint32_t>(G - 1, AS_R - 1) * 64), dim3(256), 0, s, gkey, gsum,
               gcnt, meta);
    uint64_t m[5];
    read_small(c, m, meta, 40, s);
    if (m[0]) return false;
    m[1] = m[4];   // keys in copy 0
    const uint64_t g = m[1] + (m[2] ? 1 : 0);
    PD_REQUIRE(g <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    if (!ordered) {
        launch(c, "as_emit", k_as_emit, dim3(64), dim3(256), 0, s, (const uint32_t*)gkey, (const uint64_t*)gsum,
               (const uint64_t*)gcnt, meta, okey, osum, ocnt, (uint32_t*)nullptr, cap);
        if (m[2])
            launch(c, "as_emit", k_as_emit_last, dim3(1), dim3(1), 0, s, (const uint64_t*)gsum, (const uint64_t*)meta,
                   okey, osum, ocnt, (uint32_t*)nullptr, m[1]);
    } else {
        // unordered groups with their positions, sorted by key, gathered
        Scratch st(c, g * (4 + 4 + 8 + 8 + 4), s);
        uint32_t* ek = st.as<uint32_t>();
        uint32_t* ei = ek + g;
        int64_t* es = reinterpret_cast<int64_t*>(ei + g);
        int64_t* ec = es + g;
        uint32_t* si = reinterpret_cast<uint32_t*>(ec + g);
        launch(c, "as_emit", k_as_emit, dim3(64), dim3(256), 0, s, (const uint32_t*)gkey, (const uint64_t*)gsum,
               (const uint64_t*)gcnt, meta, ek, es, ec, ei, g);
        if (m[2])
            launch(c, "as_emit", k_as_emit_last, dim3(1), dim3(1), 0, s, (const uint64_t*)gsum, (const uint64_t*)meta,
                   ek, es, ec, ei, m[1]);
        sort_pairs(c, ek, ei, 4, g, okey, si, -1, -1, s);
        launch(c, "as_gather", k_as_gather, dim3((uint32_t)std::min<uint64_t>(ceil_div(g, 256), 1024)), dim3(256), 0, s,
               g, (const uint32_t*)si, (const int64_t*)es, (const int64_t*)ec, osum, ocnt);
        PD_HIP(hipStreamSynchronize(s));   // st is released below
    }
    *groups = g;
    return true;
}

// ---------------- sorted rows: one reduce pass ----------------

constexpr uint32_t AR_BLOCK = 1024;

// Segmented (sum, count) with a flag: a run head lies in the covered rows.
struct Seg {
    uint32_t f;
    uint32_t c;
    int64_t s;
};

// x then y (x covers the earlier rows).
__device__ __forceinline__ Seg seg_join(const Seg& x, const Seg& y) {
    return y.f ? y : Seg{x.f, x.c + y.c, x.s + y.s};
}

// Inclusive segmented scan over the wave's lanes.
__device__ __forceinline__ Seg seg_wave_incl(Seg x) {
    const uint32_t l = lane();
#pragma unroll
    for (int d = 1; d < PD_WAVE; d <<= 1) {
        const uint32_t of = __shfl_up(x.f, d, PD_WAVE);
        const uint32_t oc = __shfl_up(x.c, d, PD_WAVE);
        const int64_t os = __shfl_up(x.s, d, PD_WAVE);
        if (l >= (uint32_t)d) {
            if (!x.f) {
                x.c += oc;
                x.s += os;
            }
            x.f |= of;
        }
    }
    return x;
}

// AR_ITEMS rows per thread; MINW waves per SIMD asked of the register
// allocator (8: two blocks per CU).
template <uint32_t AR_ITEMS, int MINW>
__global__ __launch_bounds__(AR_BLOCK, MINW) void k_as_runs(const uint32_t* __restrict__ key,
                                                         const int64_t* __restrict__ val, uint64_t n,
                                                         uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
                                                         uint64_t ntiles, uint32_t* __restrict__ okey,
                                                         int64_t* __restrict__ osum, int64_t* __restrict__ ocnt,
                                                         uint64_t cap, int64_t* __restrict__ cont,
                                                         uint64_t* __restrict__ total) {
    constexpr uint32_t NW = AR_BLOCK / PD_WAVE;
    constexpr uint32_t AR_TILE = AR_BLOCK * AR_ITEMS;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wh[NW];    // heads per wave, then heads before the wave
    __shared__ Seg s_ws[NW];         // segmented total per wave, then the carry into the wave
    __shared__ uint64_t s_gbase;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = lane();
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t base = tile * AR_TILE + (uint64_t)t * AR_ITEMS;
    uint32_t k[AR_ITEMS];
    int64_t v[AR_ITEMS];
    if (base + AR_ITEMS <= n) {
        // 16-byte loads (a thread's rows start 16-byte aligned)
        const uint4* kp = reinterpret_cast<const uint4*>(key + base);
#pragma unroll
        for (uint32_t j = 0; j < AR_ITEMS / 4; j++) {
            const uint4 x = kp[j];
            k[4 * j] = x.x;
            k[4 * j + 1] = x.y;
            k[4 * j + 2] = x.z;
            k[4 * j + 3] = x.w;
        }
        const longlong2* vp = reinterpret_cast<const longlong2*>(val + base);
#pragma unroll
        for (uint32_t j = 0; j < AR_ITEMS / 2; j++) {
            const longlong2 x = vp[j];
            v[2 * j] = x.x;
            v[2 * j + 1] = x.y;
        }
    } else {
#pragma unroll
        for (uint32_t j = 0; j < AR_ITEMS; j++) {
            k[j] = base + j < n ? key[base + j] : 0u;
            v[j] = base + j < n ? val[base + j] : 0;
        }
    }
    // the key before this thread's first row, and after its last
    uint32_t before = __shfl_up(k[AR_ITEMS - 1], 1, PD_WAVE);
    if (l == 0 && base > 0 && base - 1 < n) before = key[base - 1];
    uint32_t after = __shfl_down(k[0], 1, PD_WAVE);
    if (l == PD_WAVE - 1 && base + AR_ITEMS < n) after = key[base + AR_ITEMS];
    uint32_t hmask = 0;   // bit j: row j starts a run
#pragma unroll
    for (uint32_t j = 0; j < AR_ITEMS; j++) {
        const uint32_t prev = j == 0 ? before : k[j - 1];
        if (base + j < n && (base + j == 0 || k[j] != prev)) hmask |= 1u << j;
    }
    Seg a{0, 0, 0};
#pragma unroll
    for (uint32_t j = 0; j < AR_ITEMS; j++) {
        if (base + j >= n) continue;
        if (hmask >> j & 1) a = Seg{1, 1, v[j]};
        else a = Seg{a.f, a.c + 1, a.s + v[j]};
    }
    const uint32_t hc = (uint32_t)__popc(hmask);
    const uint32_t hinc = wave_incl_scan(hc);
    const Seg inc = seg_wave_incl(a);
    if (l == PD_WAVE - 1) {
        s_wh[w] = hinc;
        s_ws[w] = inc;
    }
    __syncthreads();
    if (w == 0) {
        // wave 0: what comes before each wave, the tile's head total, the
        // look-back over earlier tiles
        const uint32_t wh = l < NW ? s_wh[l] : 0u;
        const Seg ws = l < NW ? s_ws[l] : Seg{0, 0, 0};
        const uint32_t whi = wave_incl_scan(wh);
        const Seg wsi = seg_wave_incl(ws);
        const uint32_t htot = __shfl(whi, NW - 1, PD_WAVE);
        Seg wse{__shfl_up(wsi.f, 1, PD_WAVE), __shfl_up(wsi.c, 1, PD_WAVE), __shfl_up(wsi.s, 1, PD_WAVE)};
        if (l == 0) wse = Seg{0, 0, 0};
        if (l < NW) {
            s_wh[l] = whi - wh;
            s_ws[l] = wse;
        }
        const uint64_t e = lookback_single<4>(status, tile, htot);
        if (l == 0) {
            s_gbase = e;
            if (tile == ntiles - 1) *total = e + htot;
        }
    }
    // exclusive segmented value for this thread within its wave
    Seg ex{__shfl_up(inc.f, 1, PD_WAVE), __shfl_up(inc.c, 1, PD_WAVE), __shfl_up(inc.s, 1, PD_WAVE)};
    if (l == 0) ex = Seg{0, 0, 0};
    __syncthreads();
    ex = seg_join(s_ws[w], ex);
    const uint32_t hpre = s_wh[w] + hinc - hc;
    const uint64_t gbase = s_gbase;
    // walk the rows; store at each run end
    uint64_t g = gbase + hpre;   // group number + 1 of the current row
    Seg r = ex;
#pragma unroll
    for (uint32_t j = 0; j < AR_ITEMS; j++) {
        if (base + j >= n) break;
        if (hmask >> j & 1) {
            r = Seg{1, 1, v[j]};
            g++;
        } else {
            r = Seg{r.f, r.c + 1, r.s + v[j]};
        }
        const uint32_t nxt = j + 1 < AR_ITEMS ? k[j + 1] : after;
        const bool tile_end = (t == AR_BLOCK - 1 && j == AR_ITEMS - 1);
        const bool ends = base + j + 1 >= n || nxt != k[j] || tile_end;
        if (!ends) continue;
        if (r.f) {
            if (g - 1 < cap) {
                okey[g - 1] = k[j];
                osum[g - 1] = r.s;
                ocnt[g - 1] = (int64_t)r.c;
            }
        } else {
            // the run carried in from the earlier tiles (group gbase - 1)
            cont[3 * tile] = r.s;
            cont[3 * tile + 1] = (int64_t)r.c;
            cont[3 * tile + 2] = (int64_t)(gbase - 1);
        }
    }
}

__global__ __launch_bounds__(256) void k_as_fix(const int64_t* __restrict__ cont, uint64_t ntiles,
                                                int64_t* __restrict__ osum, int64_t* __restrict__ ocnt, uint64_t cap) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < ntiles; i += (uint64_t)gridDim.x * 256) {
        const int64_t c = cont[3 * i + 1];
        const uint64_t g = (uint64_t)cont[3 * i + 2];
        if (c == 0 || g >= cap) continue;
        atomicAdd((unsigned long long*)&osum[g], (unsigned long long)cont[3 * i]);
        atomicAdd((unsigned long long*)&ocnt[g], (unsigned long long)c);
    }
}

uint64_t aggr_sum_sorted(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                         int64_t* osum, int64_t* ocnt, uint64_t cap, hipStream_t s) {
    if (n == 0) return 0;
    PD_REQUIRE((uintptr_t)key % 16 == 0 && (uintptr_t)val % 16 == 0, PIMDAL_E_INTERNAL,
               "sorted rows must be 16-byte aligned");
    const int64_t items = c->tune_get("as_runs_items", 8);
    PD_REQUIRE(items == 4 || items == 8 || items == 16, PIMDAL_E_INVALID, "as_runs_items must be 4, 8 or 16");
    const uint64_t ntiles = ceil_div(n, (uint64_t)AR_BLOCK * items);
    // status[ntiles] | ticket | total | cont[3 * ntiles]
    Scratch st(c, 8 * (ntiles + 2 + 3 * ntiles), s);
    uint64_t* status = st.as<uint64_t>();
    uint64_t* total = status + ntiles + 1;
    int64_t* cont = reinterpret_cast<int64_t*>(total + 1);
    fill_u64(c, status, 0, ntiles + 2 + 3 * ntiles, s);
    uint32_t* ticket = reinterpret_cast<uint32_t*>(status + ntiles);
    if (items == 4)
        launch(c, "as_runs", k_as_runs<4, 8>, dim3((uint32_t)ntiles), dim3(AR_BLOCK), 0, s, key, val, n, status,
               ticket, ntiles, okey, osum, ocnt, cap, cont, total);
    else if (items == 8)
        launch(c, "as_runs", k_as_runs<8, 1>, dim3((uint32_t)ntiles), dim3(AR_BLOCK), 0, s, key, val, n, status,
               ticket, ntiles, okey, osum, ocnt, cap, cont, total);
    else
        launch(c, "as_runs", k_as_runs<16, 1>, dim3((uint32_t)ntiles), dim3(AR_BLOCK), 0, s, key, val, n, status,
               ticket, ntiles, okey, osum, ocnt, cap, cont, total);
    launch(c, "as_fix", k_as_fix, dim3((uint32_t)std::min<uint64_t>(ceil_div(ntiles, 256), 1024)), dim3(256), 0, s,
           (const int64_t*)cont, ntiles, osum, ocnt, cap);
    const uint64_t g = read_u64(c, total, s);
    PD_REQUIRE(g <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    return g;
}

// ---------------- dense key range: one range-partition pass ----------------
// When all keys lie within AD_MAXF * AD_SPAN (2^24, knob aggr_dense_max) of
// the smallest, partition p holds keys [kmin + p AD_SPAN, kmin + (p + 1)
// AD_SPAN).  A counting pass, a scan over tiles and one scatter move every
// row once (the 8-bit hashed passes of the general path move it twice);
// then one block per partition sums into LDS arrays indexed by the key's
// offset (no hashing, no probing) and writes the partition's groups in key
// order at its group offset, found by decoupled look-back over partitions
// taken in ticket order.  Groups leave in key order, so the hash and the
// sort aggregation both take this path.
constexpr uint32_t AD_SPAN = 8192;                  // keys per partition
constexpr uint32_t AD_SHIFT = 13;
constexpr uint32_t AD_MAXF = 2048;                  // partitions
constexpr uint32_t AD_BLOCK = 1024;
constexpr uint32_t AD_ITEMS = 8;
constexpr uint32_t AD_TILE = AD_BLOCK * AD_ITEMS;   // rows per tile
constexpr uint32_t AD_GT = 16;                      // tiles per scan group
constexpr uint32_t AD_FLAG_AGG = 1u << 30;
constexpr uint32_t AD_FLAG_PRE = 2u << 30;
constexpr uint32_t AD_VALUE = (1u << 30) - 1;

// Per-block min and max of the keys (mm[2 b], mm[2 b + 1]); the host folds
// them.  16-byte loads over the 16-byte-aligned middle.
__global__ __launch_bounds__(256) void k_ad_minmax(const uint32_t* __restrict__ key, uint64_t n,
                                                   uint32_t* __restrict__ mm) {
    __shared__ uint32_t s_lo[256 / PD_WAVE], s_hi[256 / PD_WAVE];
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t head = min(n, (uint64_t)(((16 - (uintptr_t)key % 16) % 16) / 4));
    const uint64_t nv = (n - head) / 4;
    for (uint64_t i = tid; i < head; i += stride) { lo = min(lo, key[i]); hi = max(hi, key[i]); }
    const uint4* k4 = reinterpret_cast<const uint4*>(key + head);
    uint64_t i4 = tid;
    for (; i4 + 3 * stride < nv; i4 += 4 * stride) {   // four loads in flight
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = k4[i4 + u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            lo = min(min(lo, q[u].x), min(q[u].y, min(q[u].z, q[u].w)));
            hi = max(max(hi, q[u].x), max(q[u].y, max(q[u].z, q[u].w)));
        }
    }
    for (; i4 < nv; i4 += stride) {
        const uint4 q = k4[i4];
        lo = min(min(lo, q.x), min(q.y, min(q.z, q.w)));
        hi = max(max(hi, q.x), max(q.y, max(q.z, q.w)));
    }
    for (uint64_t i = head + nv * 4 + tid; i < n; i += stride) { lo = min(lo, key[i]); hi = max(hi, key[i]); }
    for (int o = PD_WAVE / 2; o > 0; o /= 2) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    const uint32_t w = threadIdx.x / PD_WAVE;
    if (lane() == 0) { s_lo[w] = lo; s_hi[w] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t i = 1; i < 256 / PD_WAVE; i++) { lo = min(lo, s_lo[i]); hi = max(hi, s_hi[i]); }
        mm[2 * blockIdx.x] = lo;
        mm[2 * blockIdx.x + 1] = hi;
    }
}

// Min and max of a sample of the keys, rows i * stride (i < 2^16) spread
// over the whole input: per-block pairs mm[2 b], mm[2 b + 1]; 64 blocks.
constexpr uint32_t AD_SAMPLE = 1u << 16;
__global__ __launch_bounds__(256) void k_ad_sample_minmax(const uint32_t* __restrict__ key, uint64_t n,
                                                          uint64_t stride, uint32_t* __restrict__ mm) {
    __shared__ uint32_t s_lo[256 / PD_WAVE], s_hi[256 / PD_WAVE];
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    uint32_t q[4];
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) {
        const uint64_t row = ((uint64_t)(blockIdx.x * 4 + u) * 256 + threadIdx.x) * stride;
        q[u] = row < n ? key[row] : key[0];
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; u++) { lo = min(lo, q[u]); hi = max(hi, q[u]); }
    for (int o = PD_WAVE / 2; o > 0; o /= 2) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    const uint32_t w = threadIdx.x / PD_WAVE;
    if (lane() == 0) { s_lo[w] = lo; s_hi[w] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t i = 1; i < 256 / PD_WAVE; i++) { lo = min(lo, s_lo[i]); hi = max(hi, s_hi[i]); }
        mm[2 * blockIdx.x] = lo;
        mm[2 * blockIdx.x + 1] = hi;
    }
}

// Rows per partition of each AD_TILE-row tile: cnt[tile * F + p].  A key
// outside [kmin, kmin + F * AD_SPAN) is not counted and sets *outside.
// Persistent: block b counts tiles b, b + gridDim.x, ...
__global__ __launch_bounds__(512) void k_ad_count(const uint32_t* __restrict__ kin, uint64_t n, uint32_t kmin,
                                                  uint32_t F, uint32_t* __restrict__ cnt,
                                                  uint32_t* __restrict__ outside) {
    constexpr uint32_t Q = AD_TILE / 512;
    __shared__ uint32_t h[AD_MAXF];
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t ntiles = (n + AD_TILE - 1) / AD_TILE;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        for (uint32_t i = t; i < F; i += 512) h[i] = 0;
        const uint64_t wave_base = tile * AD_TILE + (uint64_t)w * (PD_WAVE * Q);
        const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
            (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * Q), n - wave_base) : 0));
        const __amdgpu_buffer_rsrc_t rk = buf_rsrc(kin + wave_base, wrows * 4u);
        uint32_t q[Q];
#pragma unroll
        for (uint32_t i = 0; i < Q; i++) q[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
        __syncthreads();
        bool out = false;
#pragma unroll
        for (uint32_t i = 0; i < Q; i++) {
            const uint32_t p = (q[i] - kmin) >> AD_SHIFT;
            if (l + i * PD_WAVE < wrows) {
                if (q[i] >= kmin && p < F) atomicAdd(&h[p], 1u);
                else out = true;
            }
        }
        if (pd::ballot(out) && l == 0) atomicOr(outside, 1u);
        __syncthreads();
        for (uint32_t i = t; i < F; i += 512) cnt[tile * F + i] = h[i];
        __syncthreads();
    }
}

// In place: cnt[tile][p] becomes the rows of partition p in earlier tiles;
// start[p] (p <= F) = rows of partitions below p.  Thread t takes
// partitions t and t + 1024; a block takes the next group of AD_GT tiles
// from a counter and resolves the group prefix by decoupled look-back.
__global__ __launch_bounds__(1024) void k_ad_tscan(uint32_t* __restrict__ cnt, uint32_t ntiles, uint32_t F,
                                                   uint32_t* __restrict__ gstatus, uint32_t* __restrict__ ctr,
                                                   uint32_t* __restrict__ start) {
    __shared__ uint32_t s_g, s_w[1024 / PD_WAVE];
    const uint32_t t = threadIdx.x, l = lane(), w = t / PD_WAVE;
    if (t == 0) s_g = atomicAdd(ctr, 1u);
    __syncthreads();
    const uint32_t g = s_g;
    const uint32_t ngroups = (ntiles + AD_GT - 1) / AD_GT;
    const uint32_t t0 = g * AD_GT, t1 = min(ntiles, t0 + AD_GT);
    uint32_t tot[2] = {0, 0};
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        const uint32_t p = t + h * 1024;
        if (p >= F) continue;
        uint32_t cg[AD_GT], sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < AD_GT; i++) cg[i] = t0 + i < t1 ? cnt[(uint64_t)(t0 + i) * F + p] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < AD_GT; i++) sum += cg[i];
        st_relaxed(&gstatus[(uint64_t)g * F + p], (g == 0 ? AD_FLAG_PRE : AD_FLAG_AGG) | sum);
        uint32_t excl = 0;
        if (g > 0) {
            int64_t j = (int64_t)g - 1;
            bool done = false;
            while (!done) {   // every earlier group is running (tickets), so each publishes
                uint32_t st[4];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    st[q] = (j - q >= 0) ? ld_relaxed(&gstatus[(uint64_t)(j - q) * F + p]) : AD_FLAG_PRE;
                bool stall = false;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (done || stall) continue;
                    if ((st[q] >> 30) == 0) { stall = true; continue; }
                    excl += st[q] & AD_VALUE;
                    if ((st[q] >> 30) == 2) done = true;
                    else j--;
                }
                if (stall) __builtin_amdgcn_s_sleep(1);
            }
            st_relaxed(&gstatus[(uint64_t)g * F + p], AD_FLAG_PRE | (excl + sum));
        }
        uint32_t run = excl;
#pragma unroll
        for (uint32_t i = 0; i < AD_GT; i++) {
            if (t0 + i < t1) cnt[(uint64_t)(t0 + i) * F + p] = run;
            run += cg[i];
        }
        tot[h] = run;
    }
    if (g == ngroups - 1) {
        // tot = rows of partitions t and t + 1024: exclusive scan over partitions
        uint32_t base = 0;
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
            const uint32_t v = tot[h];
            const uint32_t inc = wave_incl_scan(v);
            if (l == PD_WAVE - 1) s_w[w] = inc;
            __syncthreads();
            uint32_t add = 0, all = 0;
            for (uint32_t i = 0; i < 1024 / PD_WAVE; i++) {
                if (i < w) add += s_w[i];
                all += s_w[i];
            }
            const uint32_t p = t + h * 1024;
            if (p < F) start[p] = base + add + inc - v;
            if (p == F - 1) start[F] = base + add + inc;   // every row
            base += all;
            __syncthreads();
        }
    }
}

// Block b takes tile (b % 8) * xt + b / 8 (neighbouring tiles on one XCD,
// so partition runs of neighbouring tiles meet in its L2).  Rows are ranked
// within their partition by LDS counters (order inside a partition is not
// needed), staged in LDS by partition and written out in partition runs.
__global__ __launch_bounds__(AD_BLOCK) void k_ad_scatter(const uint32_t* __restrict__ kin,
                                                         const int64_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                         int64_t* __restrict__ vout, uint64_t n, uint32_t kmin,
                                                         uint32_t F, const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ start, uint32_t xt) {
    __shared__ __align__(16) uint32_t s_k[AD_TILE];
    __shared__ __align__(16) int64_t s_v[AD_TILE];
    __shared__ uint32_t s_c[AD_MAXF];   // tile rows per partition, then their start in the tile
    __shared__ uint32_t s_g[AD_MAXF];   // output position of the partition's run minus its tile start
    __shared__ uint32_t s_ws[AD_BLOCK / PD_WAVE];
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t tile = (uint64_t)(blockIdx.x % 8) * xt + blockIdx.x / 8;
    const uint64_t tile_base = tile * AD_TILE;
    if (tile_base >= n) return;   // grid rounded up to a multiple of 8
    const uint32_t tile_rows = (uint32_t)min((uint64_t)AD_TILE, n - tile_base);
    uint32_t gpos[2];
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        const uint32_t p = t + h * AD_BLOCK;
        s_c[p] = 0;
        gpos[h] = p < F ? start[p] + cnt[tile * F + p] : 0u;
    }
    const uint64_t wave_base = tile_base + (uint64_t)w * (PD_WAVE * AD_ITEMS);
    const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
        (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * AD_ITEMS), n - wave_base) : 0));
    const __amdgpu_buffer_rsrc_t rk = buf_rsrc(kin + wave_base, wrows * 4u);
    const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vin + wave_base, wrows * 8u);
    uint32_t key[AD_ITEMS];
    int64_t val[AD_ITEMS];
#pragma unroll
    for (uint32_t i = 0; i < AD_ITEMS; i++) {
        key[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0);
        val[i] = (int64_t)__builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rv, (l + i * PD_WAVE) * 8u, 0, 0));
    }
    __syncthreads();
    // rank within the partition (LDS counter), partition in the high bits
    uint32_t rk_d[AD_ITEMS];
#pragma unroll
    for (uint32_t i = 0; i < AD_ITEMS; i++) {
        rk_d[i] = 0xFFFFFFFFu;
        if (l + i * PD_WAVE < wrows) {
            const uint32_t p = (key[i] - kmin) >> AD_SHIFT;
            rk_d[i] = atomicAdd(&s_c[p], 1u) | (p << 16);
        }
    }
    __syncthreads();
    // tile-local partition starts: exclusive scan over the counters
    uint32_t base = 0;
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        const uint32_t p = t + h * AD_BLOCK;
        const uint32_t c = s_c[p];
        const uint32_t inc = wave_incl_scan(c);
        if (l == PD_WAVE - 1) s_ws[w] = inc;
        __syncthreads();
        uint32_t add = 0, all = 0;
        for (uint32_t i = 0; i < AD_BLOCK / PD_WAVE; i++) {
            if (i < w) add += s_ws[i];
            all += s_ws[i];
        }
        const uint32_t st = base + add + inc - c;
        base += all;
        __syncthreads();   // s_ws is rewritten by the next half
        s_c[p] = st;
        s_g[p] = gpos[h] - st;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < AD_ITEMS; i++) {
        if (rk_d[i] == 0xFFFFFFFFu) continue;
        const uint32_t j = s_c[rk_d[i] >> 16] + (rk_d[i] & 0xFFFFu);
        s_k[j] = key[i];
        s_v[j] = val[i];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < AD_ITEMS; i++) {
        const uint32_t j = i * AD_BLOCK + t;
        if (j >= tile_rows) break;
        const uint32_t k = s_k[j];
        const uint32_t dst = s_g[(k - kmin) >> AD_SHIFT] + j;
        kout[dst] = k;
        vout[dst] = s_v[j];
    }
}

// One block per partition, in ticket order: LDS count and sum arrays
// indexed by key - kmin - p AD_SPAN, then the groups (count > 0) in key
// order at the partition's group offset (decoupled look-back over
// partitions).  Writes at most `cap` groups; *total = all groups.
__global__ __launch_bounds__(AD_BLOCK) void k_ad_sum(const uint32_t* __restrict__ pk, const int64_t* __restrict__ pv,
                                                     const uint32_t* __restrict__ start, uint32_t F, uint32_t kmin,
                                                     uint32_t* __restrict__ ctr, uint64_t* __restrict__ status,
                                                     uint32_t* __restrict__ okey, int64_t* __restrict__ osum,
                                                     int64_t* __restrict__ ocnt, uint64_t cap,
                                                     uint64_t* __restrict__ total) {
    constexpr uint32_t SLOTS = AD_SPAN / AD_BLOCK;   // slots per thread at the emit
    __shared__ uint32_t s_cnt[AD_SPAN];
    __shared__ unsigned long long s_sum[AD_SPAN];
    __shared__ uint32_t s_p, s_w[AD_BLOCK / PD_WAVE];
    __shared__ uint64_t s_excl;
    const uint32_t t = threadIdx.x, l = lane(), w = t / PD_WAVE;
    if (t == 0) s_p = atomicAdd(ctr, 1u);
#pragma unroll
    for (uint32_t i = 0; i < SLOTS; i++) {
        s_cnt[i * AD_BLOCK + t] = 0;
        s_sum[i * AD_BLOCK + t] = 0;
    }
    __syncthreads();
    const uint32_t p = s_p;
    const uint32_t r0 = start[p], r1 = start[p + 1];
    const uint32_t kbase = kmin + p * AD_SPAN;
    // four rows per thread in flight per round
    uint32_t i = r0 + t;
    for (; i + 3 * AD_BLOCK < r1; i += 4 * AD_BLOCK) {
        uint32_t k[4];
        int64_t v[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            k[u] = pk[i + u * AD_BLOCK];
            v[u] = pv[i + u * AD_BLOCK];
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            atomicAdd(&s_cnt[k[u] - kbase], 1u);
            atomicAdd(&s_sum[k[u] - kbase], (unsigned long long)v[u]);
        }
    }
    for (; i < r1; i += AD_BLOCK) {
        const uint32_t k = pk[i];
        atomicAdd(&s_cnt[k - kbase], 1u);
        atomicAdd(&s_sum[k - kbase], (unsigned long long)pv[i]);
    }
    __syncthreads();
    // thread t owns slots [t SLOTS, (t + 1) SLOTS): groups counted, scanned
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < SLOTS; j++) c += s_cnt[t * SLOTS + j] ? 1u : 0u;
    const uint32_t inc = wave_incl_scan(c);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t add = 0, gp = 0;
    for (uint32_t j = 0; j < AD_BLOCK / PD_WAVE; j++) {
        if (j < w) add += s_w[j];
        gp += s_w[j];
    }
    if (w == 0) {
        const uint64_t ex = lookback_single<1>(status, p, gp);
        if (l == 0) s_excl = ex;
    }
    __syncthreads();
    uint64_t off = s_excl + add + inc - c;
#pragma unroll
    for (uint32_t j = 0; j < SLOTS; j++) {
        const uint32_t sl = t * SLOTS + j;
        const uint32_t cc = s_cnt[sl];
        if (!cc) continue;
        if (off < cap) {
            okey[off] = kbase + sl;
            osum[off] = (int64_t)s_sum[sl];
            ocnt[off] = cc;
        }
        off++;
    }
    if (p == F - 1 && t == 0) *total = s_excl + gp;
}

bool aggr_sum_dense(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                    int64_t* osum, int64_t* ocnt, uint64_t cap, uint64_t* groups, hipStream_t s) {
    const int64_t maxr = std::min<int64_t>(c->tune_get("aggr_dense_max", (int64_t)AD_MAXF * AD_SPAN),
                                           (int64_t)AD_MAXF * AD_SPAN);
    if (maxr <= 0 || n == 0) return false;
    // (min, max) of the keys: every key (`exact`), or a 2^16-key sample
    // spread over the input whose window the count pass then checks
    auto key_range = [&](bool exact, uint32_t* mm) {
        const uint32_t G = exact ? (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 64),
                                                                                      (uint64_t)c->cu_count * 4))
                                 : AD_SAMPLE / 1024;
        Scratch part(c, (uint64_t)G * 8, s);
        if (exact)
            launch(c, "ad_minmax", k_ad_minmax, dim3(G), dim3(256), 0, s, key, n, part.as<uint32_t>());
        else
            launch(c, "ad_minmax", k_ad_sample_minmax, dim3(G), dim3(256), 0, s, key, n,
                   std::max<uint64_t>(1, n / AD_SAMPLE), part.as<uint32_t>());
        std::vector<uint32_t> h(2 * G);
        read_small(c, h.data(), part.get(), (uint64_t)G * 8, s);
        mm[0] = 0xFFFFFFFFu;
        mm[1] = 0;
        for (uint32_t b = 0; b < G; b++) {
            mm[0] = std::min(mm[0], h[2 * b]);
            mm[1] = std::max(mm[1], h[2 * b + 1]);
        }
    };
    // knob aggr_dense_sample (default 1): the window from the sample, the
    // sample's span plus 1/64 of it and one partition either side, from 0
    // when that margin reaches it, and at most the widest window (still
    // holding the sample's keys); a key outside it sends the call to the
    // exact min/max
    const bool sampled = c->tune_get("aggr_dense_sample", 1) == 1 && n > AD_SAMPLE;
    uint32_t mm[2];
    key_range(!sampled, mm);
    const uint64_t span = (uint64_t)mm[1] - mm[0] + 1;
    if (span > (uint64_t)maxr) return false;   // a sample's span is at most the keys'
    uint64_t lo = mm[0], hi = mm[1];
    if (sampled) {
        const uint64_t margin = span / 64 + AD_SPAN;
        lo = mm[0] >= margin ? mm[0] - margin : 0;
        hi = std::min<uint64_t>(0xFFFFFFFFull, (uint64_t)mm[1] + margin);
        if (hi - lo + 1 > (uint64_t)maxr) {   // widest window, from lo where it still holds the sample's max
            if ((uint64_t)mm[1] + 1 > (uint64_t)maxr) lo = std::max<uint64_t>(lo, (uint64_t)mm[1] + 1 - (uint64_t)maxr);
            hi = std::min<uint64_t>(0xFFFFFFFFull, lo + (uint64_t)maxr - 1);
        }
    }
    Scratch flag(c, 8, s);
    for (bool check = sampled;; check = false) {
        const uint64_t range = hi - lo + 1;
        const uint32_t kmin = (uint32_t)lo;
        const uint32_t F = (uint32_t)ceil_div(range, AD_SPAN);
        const uint32_t ntiles = (uint32_t)ceil_div(n, AD_TILE);
        const uint32_t ngroups = (ntiles + AD_GT - 1) / AD_GT;
        const uint32_t xt = (ntiles + 7) / 8;
        // cnt[ntiles][F] | gstatus[ngroups][F] | start[F + 1] | 2 tickets | (8-aligned) status[F] | total
        const uint64_t w32 = (uint64_t)ntiles * F + (uint64_t)ngroups * F + F + 1 + 2;
        const uint64_t w32a = (w32 + 1) & ~1ull;
        Scratch sc(c, w32a * 4 + ((uint64_t)F + 1) * 8, s);
        uint32_t* cnt = sc.as<uint32_t>();
        uint32_t* gst = cnt + (uint64_t)ntiles * F;
        uint32_t* start = gst + (uint64_t)ngroups * F;
        uint32_t* tick = start + F + 1;
        uint64_t* pst = reinterpret_cast<uint64_t*>(cnt + w32a);
        uint64_t* tot = pst + F;
        fill_u32(c, flag.get(), 0, 1, s);
        const uint32_t cgrid = std::min<uint32_t>(ntiles, (uint32_t)c->cu_count * 4);
        launch(c, "ad_count", k_ad_count, dim3(cgrid), dim3(512), 0, s, key, n, kmin, F, cnt, flag.as<uint32_t>());
        if (check) {
            uint32_t outside = 0;
            read_small(c, &outside, flag.get(), 4, s);
            if (outside) {   // the sample missed keys beyond its window
                key_range(true, mm);
                if ((uint64_t)mm[1] - mm[0] + 1 > (uint64_t)maxr) return false;
                lo = mm[0];
                hi = mm[1];
                continue;
            }
        }
        fill_u32(c, gst, 0, (uint64_t)ngroups * F, s);
        fill_u32(c, tick, 0, 2, s);
        fill_u64(c, pst, 0, (uint64_t)F + 1, s);
        Scratch pk(c, n * 4, s), pv(c, n * 8, s);
        launch(c, "ad_tscan", k_ad_tscan, dim3(ngroups), dim3(1024), 0, s, cnt, ntiles, F, gst, tick, start);
        launch(c, "ad_scatter", k_ad_scatter, dim3(8 * xt), dim3(AD_BLOCK), 0, s, key, val, pk.as<uint32_t>(),
               pv.as<int64_t>(), n, kmin, F, (const uint32_t*)cnt, (const uint32_t*)start, xt);
        launch(c, "ad_sum", k_ad_sum, dim3(F), dim3(AD_BLOCK), 0, s, (const uint32_t*)pk.as<uint32_t>(),
               (const int64_t*)pv.as<int64_t>(), (const uint32_t*)start, F, kmin, tick + 1, pst, okey, osum, ocnt,
               cap, tot);
        const uint64_t g = read_u64(c, tot, s);
        PD_REQUIRE(g <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
        *groups = g;
        return true;
    }
}

// ---------------- direct key range: no table, no partition pass ----------------
// When a sample of the keys (the first 2^16 rows) spans at most AD_SPAN
// values, every block sums its rows into LDS arrays indexed by key - base,
// with no hashing and no probing.  The window [base, base + W) holds the
// sample's range with room either side (W a power of two, 512 <= W <=
// AD_SPAN); the LDS arrays hold AD_SPAN / W copies of it (at most one per
// wave), so few keys are not added by every wave into the same words.  The
// blocks' arrays are added into AD_DR HBM copies and the groups emitted in
// key order.  The window is worked out on the device from the sample's
// per-block minima and maxima, so the host reads back one word: the number
// of groups, or AD_NONE when the sample's span is too wide or a key fell
// outside the window (the caller then takes the table paths).
constexpr uint32_t AD_DR = 4;            // HBM copies (block b adds into copy b % AD_DR)
constexpr uint32_t AD_DU = 8;            // rows per thread per round
constexpr uint32_t AD_WMIN = 512;        // smallest window
constexpr uint32_t AD_SG = 16;           // blocks of the sample's min/max
constexpr uint64_t AD_NONE = ~0ull;

// The window from the sample's AD_SG (min, max) pairs: false when the span
// is wider than AD_SPAN.  The window stays within u32.
__device__ bool ad_window(const uint32_t* __restrict__ mm, uint32_t& base, uint32_t& wbits) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
#pragma unroll
    for (uint32_t b = 0; b < AD_SG; b++) {
        lo = min(lo, mm[2 * b]);
        hi = max(hi, mm[2 * b + 1]);
    }
    const uint64_t span = (uint64_t)hi - lo + 1;
    if (span > AD_SPAN) return false;
    wbits = 0;
    while ((1u << wbits) < AD_WMIN || ((1ull << wbits) < 2 * span && (1u << wbits) < AD_SPAN)) wbits++;
    const uint64_t W = 1ull << wbits;
    uint64_t b = lo - min((uint64_t)lo, (W - span) / 2);
    b = min(b, 0x100000000ull - W);
    base = (uint32_t)b;
    return true;
}

__global__ __launch_bounds__(AD_BLOCK) void k_ad_direct(const uint32_t* __restrict__ key,
                                                        const int64_t* __restrict__ val, uint64_t n, uint64_t chunk,
                                                        const uint32_t* __restrict__ mm,
                                                        unsigned long long* __restrict__ gcnt,
                                                        unsigned long long* __restrict__ gsum,
                                                        uint32_t* __restrict__ flag) {
    // slot AD_SPAN takes the rows outside the window (their sums are not used)
    __shared__ uint32_t s_cnt[AD_SPAN + 1];
    __shared__ unsigned long long s_sum[AD_SPAN + 1];
    const uint32_t t = threadIdx.x;
    uint32_t base = 0, wbits = 0;
    if (!ad_window(mm, base, wbits)) {
        if (blockIdx.x == 0 && t == 0) atomicOr(flag, 1u);
        return;
    }
    const uint32_t W = 1u << wbits;
    const uint32_t copies = min(AD_SPAN >> wbits, AD_BLOCK / PD_WAVE);
    const uint32_t my = ((t / PD_WAVE) % copies) << wbits;
    for (uint32_t i = t; i < AD_SPAN + 1; i += AD_BLOCK) {
        s_cnt[i] = 0;
        s_sum[i] = 0;
    }
    __syncthreads();
    const uint64_t start = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = min(n, start + chunk);
    bool out = false;
    for (uint64_t i0 = start; i0 < end; i0 += (uint64_t)AD_BLOCK * AD_DU) {
        uint32_t k[AD_DU];
        int64_t v[AD_DU];
        bool in[AD_DU];
        // loads clamped to the block's last row, so all of them are in
        // flight together
#pragma unroll
        for (uint32_t u = 0; u < AD_DU; u++) {
            const uint64_t i = i0 + (uint64_t)u * AD_BLOCK + t;
            in[u] = i < end;
            const uint64_t j = in[u] ? i : end - 1;
            k[u] = key[j] - base;
            v[u] = val[j];
        }
#pragma unroll
        for (uint32_t u = 0; u < AD_DU; u++) {
            if (!in[u]) continue;   // past the block's rows
            out |= k[u] >= W;
            const uint32_t sl = k[u] < W ? my + k[u] : AD_SPAN;
            atomicAdd(&s_cnt[sl], 1u);
            atomicAdd(&s_sum[sl], (unsigned long long)v[u]);
        }
    }
    if (__syncthreads_or(out)) {
        if (t == 0) atomicOr(flag, 1u);
        return;
    }
    const uint32_t r = blockIdx.x % AD_DR;
    for (uint32_t i = t; i < W; i += AD_BLOCK) {
        uint32_t c = 0;
        unsigned long long sm = 0;
        for (uint32_t j = 0; j < copies; j++) {
            c += s_cnt[(j << wbits) + i];
            sm += s_sum[(j << wbits) + i];
        }
        if (!c) continue;
        atomicAdd(&gcnt[(uint64_t)r * AD_SPAN + i], (unsigned long long)c);
        atomicAdd(&gsum[(uint64_t)r * AD_SPAN + i], sm);
    }
}

// The copies added up and the groups (count > 0) written in key order: one
// block, thread t owning slots [8t, 8t + 8).  Writes at most cap groups;
// *total = all groups.
__global__ __launch_bounds__(AD_BLOCK) void k_ad_direct_emit(const unsigned long long* __restrict__ gcnt,
                                                             const unsigned long long* __restrict__ gsum,
                                                             const uint32_t* __restrict__ mm,
                                                             const uint32_t* __restrict__ flag,
                                                             uint32_t* __restrict__ okey, int64_t* __restrict__ osum,
                                                             int64_t* __restrict__ ocnt, uint64_t cap,
                                                             uint64_t* __restrict__ total) {
    constexpr uint32_t SLOTS = AD_SPAN / AD_BLOCK;
    __shared__ uint32_t s_w[AD_BLOCK / PD_WAVE];
    const uint32_t t = threadIdx.x, l = lane(), w = t / PD_WAVE;
    uint32_t base = 0, wbits = 0;
    if (*flag || !ad_window(mm, base, wbits)) {
        if (t == 0) *total = AD_NONE;
        return;
    }
    uint64_t c[SLOTS], sm[SLOTS];
    uint32_t nz = 0;
#pragma unroll
    for (uint32_t j = 0; j < SLOTS; j++) {
        const uint32_t sl = t * SLOTS + j;
        c[j] = 0;
        sm[j] = 0;
#pragma unroll
        for (uint32_t r = 0; r < AD_DR; r++) {
            c[j] += gcnt[(uint64_t)r * AD_SPAN + sl];
            sm[j] += gsum[(uint64_t)r * AD_SPAN + sl];
        }
        nz += c[j] ? 1u : 0u;
    }
    const uint32_t inc = wave_incl_scan(nz);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint64_t off = inc - nz, all = 0;
    for (uint32_t i = 0; i < AD_BLOCK / PD_WAVE; i++) {
        if (i < w) off += s_w[i];
        all += s_w[i];
    }
#pragma unroll
    for (uint32_t j = 0; j < SLOTS; j++) {
        if (!c[j]) continue;
        if (off < cap) {
            okey[off] = base + t * SLOTS + j;
            osum[off] = (int64_t)sm[j];
            ocnt[off] = (int64_t)c[j];
        }
        off++;
    }
    if (t == 0) *total = all;
}

bool aggr_sum_direct(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                     int64_t* osum, int64_t* ocnt, uint64_t cap, uint64_t* groups, hipStream_t s) {
    if (n == 0 || c->tune_get("aggr_direct", 1) == 0) return false;
    // gcnt[AD_DR][AD_SPAN] | gsum[AD_DR][AD_SPAN] | total | flag | mm[2 AD_SG]
    const uint64_t words = 2ull * AD_DR * AD_SPAN + 2 + AD_SG;
    Scratch sc(c, words * 8, s);
    unsigned long long* gcnt = sc.as<unsigned long long>();
    unsigned long long* gsum = gcnt + (uint64_t)AD_DR * AD_SPAN;
    uint64_t* tot = reinterpret_cast<uint64_t*>(gsum + (uint64_t)AD_DR * AD_SPAN);
    uint32_t* flag = reinterpret_cast<uint32_t*>(tot + 1);
    uint32_t* mm = reinterpret_cast<uint32_t*>(tot + 2);
    fill_u64(c, gcnt, 0, words - AD_SG, s);
    launch(c, "ad_minmax", k_ad_minmax, dim3(AD_SG), dim3(256), 0, s, key, std::min<uint64_t>(n, 1ull << 16), mm);
    const uint32_t G = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(ceil_div(n, (uint64_t)AD_BLOCK * AD_DU), (uint64_t)c->cu_count));
    const uint64_t chunk = ceil_div(n, G);
    launch(c, "ad_direct", k_ad_direct, dim3(G), dim3(AD_BLOCK), 0, s, key, val, n, chunk, (const uint32_t*)mm, gcnt,
           gsum, flag);
    launch(c, "ad_direct", k_ad_direct_emit, dim3(1), dim3(AD_BLOCK), 0, s, (const unsigned long long*)gcnt,
           (const unsigned long long*)gsum, (const uint32_t*)mm, (const uint32_t*)flag, okey, osum, ocnt, cap, tot);
    const uint64_t g = read_u64(c, tot, s);
    if (g == AD_NONE) return false;
    PD_REQUIRE(g <= cap, PIMDAL_E_CAPACITY, "group-by output larger than capacity");
    *groups = g;
    return true;
}

}  // namespace pd
