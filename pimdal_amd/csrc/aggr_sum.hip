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

}  // namespace pd
