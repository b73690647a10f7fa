// serve.hip — serving loop over device-resident tables.
//
// Two ways to run a stream of queries over the resident fact and dimension
// tables:
//  * mode PIMDAL_SERVE_SEQUENTIAL — one query at a time through the plain
//    operator entry points (the drop-in path): select_range over the price
//    column; for a join, a select of the qualifying dimension rows followed
//    by hash_join against the fact slice (a select over the resident join
//    index when there is one); the fused Q1 group-by over the slice.  Each query completes before the next starts.
//  * mode PIMDAL_SERVE_BATCHED — continuous batching: arrivals queue up
//    (one FIFO per tenant) while earlier batches run; whenever one of
//    `slots` stream slots is free, arrived queries are coalesced into one
//    batch, taken round-robin over the tenants, as long as the batch's
//    block of the serving heap (the footprint of its queries' outputs,
//    statuses and partials) fits the heap's largest free range and each
//    tenant stays within its share of the heap.  A batch is one hipGraph replay per slot:
//    descriptor copy -> reset -> persistent select / join / group-by kernels
//    that walk a device work list of (query, tile) items -> group-by
//    finalisation -> result copy.  Joins read a foreign-key join index (the
//    dimension row and date of every fact row, built once at server
//    creation from a dimension-key index, since both tables stay resident),
//    or with serve_join_index = 0 probe the dimension-key index per row.
//    Slots run on their own streams, so batches overlap on the GPU.  The
//    sequential mode's joins read the same join index (one select over it),
//    so the two modes differ only in batching.
// Both modes report per query the result count and an order-independent
// checksum (include/pimdal_gen.h), so results can be compared across modes
// and with the CPU restatement.
#include <string.h>
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"
#include "../../include/pimdal_gen.h"

namespace pd {

constexpr int SV_MAXQ = 512;          // queries per kind per batch
constexpr int SV_BLOCK = 256;
constexpr int SV_TILE = 4096;         // rows per select / join work item
constexpr uint32_t SV_GB_CHUNK = 1u << 16;  // rows per group-by work item
constexpr int SV_G = 8;               // dense group codes in served group-bys

struct SvQuery {
    int64_t param;
    uint64_t offset, rows;
    uint64_t item0;      // first work item of this query within its kind
    uint64_t out_off;    // output offset in the arena (elements of the kind's output type)
    uint64_t status_off; // first look-back word of this query
    uint32_t qid;        // index in the batch's result array
    uint32_t pad;
};

struct SvKind {
    uint32_t nq, pad;
    uint64_t items;      // total work items of this kind
    SvQuery q[SV_MAXQ];
};

// One batch: its queries by kind, and the byte offsets of its block's
// regions in the serving heap (select outputs, join outputs, look-back
// status words, group-by partials), set at admission.
struct SvBatch {
    SvKind sel, join, gb;
    uint64_t status_words;
    uint64_t arena_off, status_off, partials_off;
};

struct SvResult {
    uint64_t count, checksum;
};

struct SvTables {
    const int32_t* orderkey;
    const int64_t* qty;
    const int64_t* price;
    const int64_t* disc;
    const int64_t* tax;
    const int32_t* date;
    const uint8_t* flag1;
    const uint8_t* flag2;
    uint64_t n_fact;
    const int32_t* dim_date;
    const uint4* dim_index;     // (key, dim row + 1, dim date, 0); row 0 = empty slot
    uint64_t dim_mask;
    const uint8_t* lut;         // flag codes: [0,256) flag1, [256,512) flag2
    uint32_t n2;
    // foreign-key join index over the resident tables (NULL when off):
    // dimension row of each fact row (0xFFFFFFFF = no match) and its date
    const uint32_t* fk_row;
    const int32_t* fk_date;
};

__device__ __forceinline__ uint32_t sv_find_query(const SvKind& k, uint64_t wi) {
    uint32_t lo = 0, hi = k.nq;  // last q with item0 <= wi
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (k.q[mid].item0 <= wi) lo = mid; else hi = mid;
    }
    return lo;
}

// Index slots are 16 bytes (key, dim row + 1, dim date, 0), row 0 = empty,
// linear probing at load <= 0.5: one slot read per probe, and the date comes
// with the slot, so the join predicate needs no second random read.

__global__ void k_sv_reset(const SvBatch* __restrict__ b, uint8_t* __restrict__ heap, SvResult* __restrict__ res,
                           uint32_t nres, uint32_t* __restrict__ ctr) {
    const uint64_t n = b->status_words;
    uint64_t* status = reinterpret_cast<uint64_t*>(heap + b->status_off);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        status[i] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nres; i += gridDim.x * blockDim.x) res[i] = SvResult{0, 0};
    if (blockIdx.x == 0 && threadIdx.x < 4) ctr[threadIdx.x] = 0;
}

// Persistent select / join over (query, 4096-row tile) work items.  Per
// query: decoupled look-back over its own tiles (items are taken in order,
// so a predecessor tile is always held by a running block), order-
// preserving output, checksum added atomically (a sum, so the order of the
// additions does not matter; a select's terms carry each row's output
// position, so its checksum does depend on the output order).
template <bool JOIN>
__global__ __launch_bounds__(SV_BLOCK) void k_sv_scan(const SvBatch* __restrict__ b, SvTables T,
                                                      uint8_t* __restrict__ heap, SvResult* __restrict__ res,
                                                      uint32_t* __restrict__ ctr) {
    void* arena = heap + b->arena_off;
    uint64_t* status = reinterpret_cast<uint64_t*>(heap + b->status_off);
    __shared__ uint64_t s_wi;
    __shared__ uint32_t s_wave_tot[SV_BLOCK / PD_WAVE];
    __shared__ uint64_t s_excl;
    const SvKind& K = JOIN ? b->join : b->sel;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    constexpr int RPT = SV_TILE / SV_BLOCK;  // 16 rows per thread: row = tile_base + i*256 + t
    uint32_t ck_qid = 0xFFFFFFFFu;   // the wave's checksum, flushed when its query changes
    uint64_t ck_acc = 0;
    while (true) {
        if (t == 0) s_wi = atomicAdd(&ctr[JOIN ? 1 : 0], 1u);
        __syncthreads();
        const uint64_t wi = s_wi;
        if (wi >= K.items) {
            if (l == 0 && ck_acc) atomicAdd((unsigned long long*)&res[ck_qid].checksum, (unsigned long long)ck_acc);
            break;
        }
        const uint32_t qi = sv_find_query(K, wi);
        const SvQuery& q = K.q[qi];
        const uint64_t tile = wi - q.item0;
        const uint64_t ntl = (q.rows + SV_TILE - 1) / SV_TILE;
        const uint64_t r0 = q.offset + tile * SV_TILE;
        const uint64_t rend = q.offset + q.rows;
        uint32_t hit = 0;
        uint32_t dimrow[JOIN ? RPT : 1];
        if constexpr (JOIN) {
          if (T.fk_row) {
            // resident join index: the match and its date per fact row
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                // loads clamped to the query's last row (r0 < rend): none
                // waits behind a branch
                const uint64_t r = r0 + (uint64_t)i * SV_BLOCK + t;
                const uint64_t rc = r < rend ? r : rend - 1;
                const uint32_t dr = T.fk_row[rc];
                const int32_t dd = T.fk_date[rc];
                dimrow[i] = r < rend ? dr : 0xFFFFFFFFu;
                const bool m = dimrow[i] != 0xFFFFFFFFu && (int64_t)dd < q.param;
                hit |= (uint32_t)m << i;
            }
          } else {
            // all keys, then all first-slot probes in flight together; the
            // rare longer chains are walked afterwards
            int32_t key[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const uint64_t r = r0 + (uint64_t)i * SV_BLOCK + t;
                key[i] = T.orderkey[r < rend ? r : rend - 1];   // clamped (r0 < rend)
            }
            uint4 e[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++) e[i] = T.dim_index[pd_mix64((uint32_t)key[i]) & T.dim_mask];
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const uint64_t r = r0 + (uint64_t)i * SV_BLOCK + t;
                bool m = false;
                dimrow[i] = 0xFFFFFFFFu;
                if (r < rend) {
                    uint4 x = e[i];
                    if (x.y != 0 && x.x != (uint32_t)key[i]) {
                        uint64_t hh = pd_mix64((uint32_t)key[i]) & T.dim_mask;
                        do {
                            hh = (hh + 1) & T.dim_mask;
                            x = T.dim_index[hh];
                        } while (x.y != 0 && x.x != (uint32_t)key[i]);
                    }
                    if (x.y != 0) {
                        dimrow[i] = x.y - 1;
                        m = (int64_t)(int32_t)x.z < q.param;
                    }
                }
                hit |= (uint32_t)m << i;
            }
          }
        } else {
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const uint64_t r = r0 + (uint64_t)i * SV_BLOCK + t;
                const bool m = r < rend && T.price[r < rend ? r : rend - 1] < q.param;   // clamped (r0 < rend)
                hit |= (uint32_t)m << i;
            }
        }
        // order (i, wave, lane): per item i, counts per wave
        __shared__ uint32_t s_cnt[RPT][SV_BLOCK / PD_WAVE];
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const bool m = (hit >> i) & 1;
            const uint64_t bal = pd::ballot(m);
            if (l == 0) s_cnt[i][w] = pd::popc64(bal);
        }
        __syncthreads();
        if (w == 0) {
            const uint32_t v = (&s_cnt[0][0])[l];
            const uint32_t inc = pd::wave_incl_scan(v);
            (&s_cnt[0][0])[l] = inc - v;
            if (l == PD_WAVE - 1) s_wave_tot[0] = inc;
        }
        __syncthreads();
        const uint32_t total = s_wave_tot[0];
        if (w == 0) {
            const uint64_t e = pd::lookback_single<4>(status + q.status_off, tile, total);
            if (l == 0) {
                s_excl = e;
                if (tile == ntl - 1) res[q.qid].count = e + total;
            }
        }
        __syncthreads();
        const uint64_t excl = s_excl;
        const uint64_t lt = pd::lanemask_lt();
        uint64_t ck = 0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const bool m = (hit >> i) & 1;
            const uint64_t bal = pd::ballot(m);
            if (m) {
                const uint64_t o = excl + s_cnt[i][w] + pd::popc64(bal & lt);
                const uint64_t r = r0 + (uint64_t)i * SV_BLOCK + t;
                if (o < q.rows) {
                    if constexpr (JOIN) {
                        reinterpret_cast<uint2*>(arena)[q.out_off + o] = make_uint2((uint32_t)r, dimrow[i]);
                        ck += pd_ck_join(r, dimrow[i]);
                    } else {
                        reinterpret_cast<uint32_t*>(arena)[q.out_off + o] = (uint32_t)r;
                        ck += pd_ck_select(r, o);
                    }
                }
            }
        }
        ck = pd::wave_sum64(ck);
        if (q.qid != ck_qid) {
            if (l == 0 && ck_acc) atomicAdd((unsigned long long*)&res[ck_qid].checksum, (unsigned long long)ck_acc);
            ck_acc = 0;
            ck_qid = q.qid;
        }
        ck_acc += ck;
        __syncthreads();
    }
}

// ---- streaming select and index join: 16384-row tiles, 1024-thread blocks ----
// The persistent scan above takes 4096-row tiles in 256-thread blocks and
// waits on each tile's look-back before its next loads; measured on a
// 2000-query stream it read 0.64 TB/s.  Here each block takes a tile
// (wave w: rows wbase .. wbase + 1023, lane l: rows g*256 + 4l .. +3),
// stages its kept rows compacted in the wave's LDS slice, publishes the
// tile's count, then loads its next tile while the look-back of the held
// one resolves, and writes it out (the pipeline of the library's one-pass
// select).  Tiles of a query start at its offset rounded down to 4 rows,
// so every 16-byte load is aligned; rows outside the slice are masked.
// Select: price < param, output the fact row.  Join (resident join index):
// dimension row present and its date < param, output (fact row, dim row).
constexpr int SV2_BLOCK = 1024;
constexpr int SV2_WAVES = SV2_BLOCK / PD_WAVE;
constexpr int SV2_WROWS = 1024;
constexpr uint64_t SV2_TILE = (uint64_t)SV2_WAVES * SV2_WROWS;

__host__ __device__ inline uint64_t sv2_tiles(uint64_t offset, uint64_t rows) {
    return ((offset & 3ull) + rows + SV2_TILE - 1) / SV2_TILE;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t sv2_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <bool JOIN>
__global__ __launch_bounds__(SV2_BLOCK) void k_sv_stream(const SvBatch* __restrict__ b, SvTables T,
                                                         uint8_t* __restrict__ heap, SvResult* __restrict__ res,
                                                         uint32_t* __restrict__ ctr) {
    const SvKind& K = JOIN ? b->join : b->sel;
    uint64_t* status = reinterpret_cast<uint64_t*>(heap + b->status_off);
    uint32_t* out32 = reinterpret_cast<uint32_t*>(heap + b->arena_off);
    uint2* out64 = reinterpret_cast<uint2*>(heap + b->arena_off);
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_wave_tot[SV2_WAVES];
    __shared__ uint64_t s_excl;
    __shared__ uint16_t s_idx[SV2_WAVES][SV2_WROWS];
    __shared__ uint32_t s_pay[JOIN ? SV2_WAVES : 1][JOIN ? SV2_WROWS : 1];
    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint32_t nitems = (uint32_t)K.items;

    // wave geometry of work item wi: query, tile, first row (absolute) and
    // the range [vlo, vhi) of the wave's rows inside the query's slice
    struct Geo { uint32_t qi; uint32_t tile; uint64_t wbase; uint32_t vlo, vhi; };
    auto geo = [&](uint32_t wi) -> Geo {
        Geo g;
        g.qi = sv_find_query(K, wi);
        const SvQuery& q = K.q[g.qi];
        g.tile = (uint32_t)(wi - q.item0);
        g.wbase = (q.offset & ~3ull) + (uint64_t)g.tile * SV2_TILE + (uint64_t)w * SV2_WROWS;
        const uint64_t end = q.offset + q.rows;
        g.vlo = g.wbase >= q.offset ? 0u : (uint32_t)min<uint64_t>(SV2_WROWS, q.offset - g.wbase);
        g.vhi = end <= g.wbase ? 0u : (uint32_t)min<uint64_t>(SV2_WROWS, end - g.wbase);
        return g;
    };
    // per lane: rows g*256 + 4l + e of the wave (g, e < 4)
    int64_t kp[4][4];     // select: price
    uint32_t kr[4][4];    // join: dimension row
    int32_t kd[4][4];     // join: dimension date
    auto load = [&](const Geo& g) {
        const uint32_t l4 = l * 4;
        if constexpr (JOIN) {
            const __amdgpu_buffer_rsrc_t rr = sv2_rsrc(T.fk_row + g.wbase, g.vhi * 4u);
            const __amdgpu_buffer_rsrc_t rd = sv2_rsrc(T.fk_date + g.wbase, g.vhi * 4u);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t r0 = q * 256 + l4;
                const uint4 a = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rr, r0 * 4, 0, 0));
                const uint4 d = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rd, r0 * 4, 0, 0));
                kr[q][0] = a.x; kr[q][1] = a.y; kr[q][2] = a.z; kr[q][3] = a.w;
                kd[q][0] = (int32_t)d.x; kd[q][1] = (int32_t)d.y; kd[q][2] = (int32_t)d.z; kd[q][3] = (int32_t)d.w;
            }
        } else {
            const __amdgpu_buffer_rsrc_t rp = sv2_rsrc(T.price + g.wbase, g.vhi * 8u);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t r0 = q * 256 + l4;
                const uint4 a = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rp, r0 * 8, 0, 0));
                const uint4 c = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rp, r0 * 8 + 16, 0, 0));
                kp[q][0] = (int64_t)(((uint64_t)a.y << 32) | a.x); kp[q][1] = (int64_t)(((uint64_t)a.w << 32) | a.z);
                kp[q][2] = (int64_t)(((uint64_t)c.y << 32) | c.x); kp[q][3] = (int64_t)(((uint64_t)c.w << 32) | c.z);
            }
        }
    };
    // predicate, wave ranks (packed 16-bit group counts, one 64-bit scan),
    // kept rows staged compacted in the wave's slice; returns the count
    auto stage = [&](const Geo& g) -> uint32_t {
        const int64_t param = K.q[g.qi].param;
        const uint32_t l4 = l * 4;
        uint32_t bits = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t r = q * 256 + l4 + e;
                bool keep = r >= g.vlo && r < g.vhi;
                if constexpr (JOIN) keep = keep && kr[q][e] != 0xFFFFFFFFu && (int64_t)kd[q][e] < param;
                else keep = keep && kp[q][e] < param;
                bits |= (uint32_t)keep << (q * 4 + e);
            }
        uint64_t packed = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) packed |= (uint64_t)__popc((bits >> (q * 4)) & 0xF) << (16 * q);
        const uint64_t incl = pd::wave_incl_scan64(packed);
        const uint64_t tot = __shfl(incl, PD_WAVE - 1, PD_WAVE);
        uint32_t wave_total = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t gb = (bits >> (q * 4)) & 0xF;
            uint32_t pos = wave_total + ((uint32_t)(incl >> (16 * q)) & 0xFFFF) - __popc(gb);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (gb & (1u << e)) {
                    s_idx[w][pos] = (uint16_t)(q * 256 + l4 + e);
                    if constexpr (JOIN) s_pay[w][pos] = kr[q][e];
                    pos++;
                }
            }
            wave_total += (uint32_t)(tot >> (16 * q)) & 0xFFFF;
        }
        return wave_total;
    };
    auto block_totals = [&](uint32_t& woff, uint32_t& ttot) {
        woff = 0;
        ttot = 0;
#pragma unroll
        for (int i = 0; i < SV2_WAVES; i++) {
            const uint32_t v = s_wave_tot[i];
            woff += (i < (int)w) ? v : 0;
            ttot += v;
        }
    };

    // prologue: the first tile counted, staged and published
    if (t == 0) s_item = atomicAdd(&ctr[JOIN ? 1 : 0], 1u);
    __syncthreads();
    uint32_t ia = __builtin_amdgcn_readfirstlane(s_item);
    if (ia >= nitems) return;
    Geo A = geo(ia);
    load(A);
    uint32_t wt = stage(A);
    if (l == 0) s_wave_tot[w] = wt;
    __syncthreads();
    uint32_t woff, ttot;
    block_totals(woff, ttot);
    if (w == 0) pd::lb_publish_agg(status + K.q[A.qi].status_off, A.tile, ttot);
    // the wave's checksum of its current query, added to the query's word
    // only when the wave moves to another query (same-word device atomics
    // serialize: one per wave per tile would queue thousands on a query)
    uint32_t ck_qid = 0xFFFFFFFFu;
    uint64_t ck_acc = 0;
    while (true) {
        __syncthreads();   // every wave has read s_item and s_wave_tot
        if (t == 0) s_item = atomicAdd(&ctr[JOIN ? 1 : 0], 1u);
        __syncthreads();
        const uint32_t ib = __builtin_amdgcn_readfirstlane(s_item);
        Geo B{};
        if (ib < nitems) {
            B = geo(ib);
            load(B);
        }
        const SvQuery& qa = K.q[A.qi];
        if (w == 0) {
            const uint64_t excl = pd::lb_resolve<1>(status + qa.status_off, A.tile, ttot);
            if (l == 0) {
                s_excl = excl;
                if ((uint64_t)A.tile + 1 == sv2_tiles(qa.offset, qa.rows)) res[qa.qid].count = excl + ttot;
            }
        }
        __syncthreads();
        {
            const uint64_t obase = s_excl + woff;
            uint64_t ck = 0;
            for (uint32_t i = l; i < wt; i += PD_WAVE) {
                const uint64_t o = obase + i;
                const uint32_t row = (uint32_t)(A.wbase + s_idx[w][i]);
                if constexpr (JOIN) {
                    const uint32_t dr = s_pay[w][i];
                    out64[qa.out_off + o] = make_uint2(row, dr);
                    ck += pd_ck_join(row, dr);
                } else {
                    out32[qa.out_off + o] = row;
                    ck += pd_ck_select(row, o);
                }
            }
            ck = pd::wave_sum64(ck);
            if (qa.qid != ck_qid) {
                if (l == 0 && ck_acc) atomicAdd((unsigned long long*)&res[ck_qid].checksum, (unsigned long long)ck_acc);
                ck_acc = 0;
                ck_qid = qa.qid;
            }
            ck_acc += ck;
        }
        if (ib >= nitems) {
            if (l == 0 && ck_acc) atomicAdd((unsigned long long*)&res[ck_qid].checksum, (unsigned long long)ck_acc);
            break;
        }
        // the wave's own LDS slice is free again (its reads were issued first)
        wt = stage(B);
        if (l == 0) s_wave_tot[w] = wt;
        __syncthreads();
        block_totals(woff, ttot);
        A = B;
        if (w == 0) pd::lb_publish_agg(status + K.q[A.qi].status_off, A.tile, ttot);
    }
}

// Persistent dense group-by over (query, 2^16-row chunk) items: Q1 measures
// in registers for up to 8 flag codes, one partial row per item.
__global__ __launch_bounds__(SV_BLOCK) void k_sv_groupby(const SvBatch* __restrict__ b, SvTables T,
                                                         uint8_t* __restrict__ heap, uint32_t* __restrict__ ctr) {
    uint64_t* partials = reinterpret_cast<uint64_t*>(heap + b->partials_off);
    __shared__ uint64_t s_wi;
    __shared__ uint64_t s_red[SV_BLOCK / PD_WAVE][SV_G][8];
    const SvKind& K = b->gb;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    while (true) {
        if (t == 0) s_wi = atomicAdd(&ctr[2], 1u);
        __syncthreads();
        const uint64_t wi = s_wi;
        if (wi >= K.items) break;
        const SvQuery& q = K.q[sv_find_query(K, wi)];
        const uint64_t c0 = q.offset + (wi - q.item0) * SV_GB_CHUNK;
        const uint64_t c1 = min(q.offset + q.rows, c0 + SV_GB_CHUNK);
        int64_t aq[SV_G], ap[SV_G], ad[SV_G], adp[SV_G], an[SV_G];
        unsigned __int128 ach[SV_G];
#pragma unroll
        for (int g = 0; g < SV_G; g++) { aq[g] = ap[g] = ad[g] = adp[g] = an[g] = 0; ach[g] = 0; }
        auto add_row = [&](int32_t dt, uint32_t a, uint32_t b, int64_t qq, int64_t p, int64_t d, int64_t x) {
            if ((int64_t)dt > q.param) return;
            const uint32_t code = (uint32_t)T.lut[a] * T.n2 + T.lut[256 + b];
            const int64_t dp = p * (100 - d);
            const __int128 ch = (__int128)dp * (100 + x);
#pragma unroll
            for (int g = 0; g < SV_G; g++) {
                const bool on = code == (uint32_t)g;
                aq[g] += on ? qq : 0;
                ap[g] += on ? p : 0;
                ad[g] += on ? d : 0;
                adp[g] += on ? dp : 0;
                ach[g] += on ? (unsigned __int128)ch : 0;
                an[g] += on ? 1 : 0;
            }
        };
        auto scalar = [&](uint64_t r) {
            add_row(T.date[r], T.flag1[r], T.flag2[r], T.qty[r], T.price[r], T.disc[r], T.tax[r]);
        };
        // rows [c0, a0) and [a1, c1) one at a time; the 4-aligned middle four
        // rows per thread per step with 16-byte loads (the columns are
        // 256-byte aligned allocations)
        const uint64_t a0 = min(c1, (c0 + 3) & ~3ull), a1 = max(a0, c1 & ~3ull);
        for (uint64_t r = c0 + t; r < a0; r += SV_BLOCK) scalar(r);
        for (uint64_t v = a0 / 4 + t; v < a1 / 4; v += SV_BLOCK) {
            const int4 dt = reinterpret_cast<const int4*>(T.date)[v];
            const uint32_t fa = reinterpret_cast<const uint32_t*>(T.flag1)[v];
            const uint32_t fb = reinterpret_cast<const uint32_t*>(T.flag2)[v];
            const longlong2 q0 = reinterpret_cast<const longlong2*>(T.qty)[2 * v], q1 = reinterpret_cast<const longlong2*>(T.qty)[2 * v + 1];
            const longlong2 p0 = reinterpret_cast<const longlong2*>(T.price)[2 * v], p1 = reinterpret_cast<const longlong2*>(T.price)[2 * v + 1];
            const longlong2 d0 = reinterpret_cast<const longlong2*>(T.disc)[2 * v], d1 = reinterpret_cast<const longlong2*>(T.disc)[2 * v + 1];
            const longlong2 x0 = reinterpret_cast<const longlong2*>(T.tax)[2 * v], x1 = reinterpret_cast<const longlong2*>(T.tax)[2 * v + 1];
            add_row(dt.x, fa & 0xFF, fb & 0xFF, q0.x, p0.x, d0.x, x0.x);
            add_row(dt.y, (fa >> 8) & 0xFF, (fb >> 8) & 0xFF, q0.y, p0.y, d0.y, x0.y);
            add_row(dt.z, (fa >> 16) & 0xFF, (fb >> 16) & 0xFF, q1.x, p1.x, d1.x, x1.x);
            add_row(dt.w, fa >> 24, fb >> 24, q1.y, p1.y, d1.y, x1.y);
        }
        for (uint64_t r = a1 + t; r < c1; r += SV_BLOCK) scalar(r);
#pragma unroll
        for (int g = 0; g < SV_G; g++) {
            const uint64_t lo = (uint64_t)ach[g];
            uint64_t v[8] = {(uint64_t)aq[g], (uint64_t)ap[g], (uint64_t)ad[g], (uint64_t)adp[g],
                             lo & 0xFFFFFFFFull, (uint64_t)(ach[g] >> 64), (uint64_t)an[g], lo >> 32};
#pragma unroll
            for (int f = 0; f < 8; f++) v[f] = pd::wave_sum64(v[f]);
            if (l == 0)
#pragma unroll
                for (int f = 0; f < 8; f++) s_red[w][g][f] = v[f];
        }
        __syncthreads();
        if (t < SV_G * 8) {
            const uint32_t g = t / 8, f = t % 8;
            uint64_t s = 0;
            for (int ww = 0; ww < SV_BLOCK / PD_WAVE; ww++) s += s_red[ww][g][f];
            partials[(wi * SV_G + g) * 8 + f] = s;
        }
        __syncthreads();
    }
}

// One 512-thread block per group-by query, wave g = group code g: the
// lanes sum the query's item partials (lane l: items l, l + 64, ...), then
// the wave.  Every field is a 64-bit wrapping sum: the int64 measures keep
// their low 64 bits, the count and the two 32-bit halves of the charge's
// low word stay below 2^58 (2^16 rows x 2^32 per item, at most 2^10 items
// below 2^31 rows), and the charge's high word enters the 128-bit total
// shifted by 64, so only its low 64 bits count.  Then count and checksum of
// the non-empty groups (keys flag1*256+flag2).
__global__ __launch_bounds__(SV_G * PD_WAVE) void k_sv_groupby_final(const SvBatch* __restrict__ b,
                                                                     const uint8_t* __restrict__ heap,
                                                                     const uint8_t* __restrict__ inv, uint32_t n1,
                                                                     uint32_t n2, SvResult* __restrict__ res) {
    const uint64_t* partials = reinterpret_cast<const uint64_t*>(heap + b->partials_off);
    const SvKind& K = b->gb;
    if (blockIdx.x >= K.nq) return;
    const SvQuery& q = K.q[blockIdx.x];
    const uint64_t items = (q.rows + SV_GB_CHUNK - 1) / SV_GB_CHUNK;
    __shared__ uint64_t s_ck[SV_G];
    __shared__ uint32_t s_n[SV_G];
    const uint32_t g = threadIdx.x / PD_WAVE, l = pd::lane();
    uint64_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (g < n1 * n2) {
        for (uint64_t it = l; it < items; it += PD_WAVE) {
            const uint64_t* r = partials + ((q.item0 + it) * SV_G + g) * 8;
#pragma unroll
            for (int f = 0; f < 8; f++) a[f] += r[f];
        }
    }
#pragma unroll
    for (int f = 0; f < 8; f++) a[f] = pd::wave_sum64(a[f]);
    if (l == 0) {
        s_ck[g] = 0;
        s_n[g] = 0;
        if (g < n1 * n2 && a[6]) {
            const unsigned __int128 ch = (unsigned __int128)a[4] + ((unsigned __int128)a[7] << 32) +
                                         ((unsigned __int128)a[5] << 64);
            const int64_t v[7] = {(int64_t)a[0], (int64_t)a[1], (int64_t)a[2], (int64_t)a[3],
                                  (int64_t)(uint64_t)ch, (int64_t)(uint64_t)(ch >> 64), (int64_t)a[6]};
            const int64_t key = (int64_t)inv[g / n2] * 256 + inv[256 + g % n2];
            s_ck[g] = pd_ck_group(key, v);
            s_n[g] = 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ck = 0, n = 0;
        for (int i = 0; i < SV_G; i++) { ck += s_ck[i]; n += s_n[i]; }
        res[q.qid].count = n;
        res[q.qid].checksum = ck;
    }
}

__global__ void k_dim_index_build(const int32_t* __restrict__ key, const int32_t* __restrict__ date, uint64_t n,
                                  uint4* __restrict__ idx, uint64_t mask, uint32_t* __restrict__ dup) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        // claim a slot with its first 8 bytes (key, row + 1), then add the date
        const uint64_t want = ((i + 1) << 32) | (uint32_t)key[i];
        uint64_t h = pd_mix64((uint32_t)key[i]) & mask;
        while (true) {
            unsigned long long* slot = reinterpret_cast<unsigned long long*>(&idx[h]);
            const uint64_t prev = atomicCAS(slot, 0ull, (unsigned long long)want);
            if (prev == 0) {
                idx[h].z = (uint32_t)date[i];
                break;
            }
            if ((uint32_t)prev == (uint32_t)key[i]) { atomicOr(dup, 1u); break; }
            h = (h + 1) & mask;
        }
    }
}

// Foreign-key join index: for every fact row, its dimension row (probing
// the key index once, at server creation) and that row's date.
__global__ void k_fk_index_build(SvTables T, uint32_t* __restrict__ fk_row, int32_t* __restrict__ fk_date) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < T.n_fact; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t key = (uint32_t)T.orderkey[r];
        uint64_t h = pd_mix64(key) & T.dim_mask;
        uint4 x = T.dim_index[h];
        while (x.y != 0 && x.x != key) {
            h = (h + 1) & T.dim_mask;
            x = T.dim_index[h];
        }
        fk_row[r] = x.y != 0 ? x.y - 1 : 0xFFFFFFFFu;
        // no dimension row: the largest date, so no date predicate keeps the row
        fk_date[r] = x.y != 0 ? (int32_t)x.z : INT32_MAX;
    }
}

// checksum of select / join outputs of the sequential path
__global__ void k_ck_rows(const uint32_t* __restrict__ rows, uint64_t n, uint64_t base, uint64_t* __restrict__ out) {
    uint64_t s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += pd_ck_select(base + rows[i], i);
    s = pd::wave_sum64(s);
    if (pd::lane() == 0 && s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

__global__ void k_ck_pairs(const uint32_t* __restrict__ fact, const uint32_t* __restrict__ dim, uint64_t n,
                           uint64_t fact_base, uint64_t* __restrict__ out) {
    uint64_t s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += pd_ck_join(fact_base + fact[i], dim[i]);
    s = pd::wave_sum64(s);
    if (pd::lane() == 0 && s) atomicAdd((unsigned long long*)out, (unsigned long long)s);
}

__global__ void k_iota64(uint32_t* __restrict__ out, uint64_t n, uint64_t start) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)(start + i);
}

}  // namespace pd

using namespace pd;

// First-fit allocator over the serving heap's byte range, 256-byte granules,
// neighbouring free ranges merged.  Host-side bookkeeping only: the memory
// is one pool allocation made when the server is created.
class SvHeap {
public:
    void init(uint64_t bytes) {
        free_.clear();
        size_ = bytes / 256 * 256;
        if (size_) free_[0] = size_;
        used_ = peak_ = 0;
    }
    static uint64_t round(uint64_t n) { return (std::max<uint64_t>(n, 1) + 255) / 256 * 256; }
    uint64_t largest() const {
        uint64_t m = 0;
        for (const auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    bool alloc(uint64_t n, uint64_t* off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n;
            const uint64_t at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = round(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t size() const { return size_; }
    uint64_t peak() const { return peak_; }
    void reset_peak() { peak_ = used_; }
private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

struct pimdal_server {
    pimdal_ctx* ctx = nullptr;
    pimdal_tables tab{};
    SvTables T{};
    void* dim_index = nullptr;
    void* fk_row = nullptr;   // foreign-key join index (optional)
    void* fk_date = nullptr;
    void* lut = nullptr;      // device [512] codes + [512] inverse
    uint8_t h_inv[512];
    uint32_t n1 = 0, n2 = 0;
    int nslots = 0;
    uint8_t* heap = nullptr;  // serving heap: every batch's outputs, statuses and partials
    SvHeap heap_alloc;
    // streaming kernels (16384-row tiles) for selects and index joins;
    // otherwise the 4096-row scan (unaligned columns, probe-index joins)
    bool stream_sel = false, stream_join = false;
    uint32_t stream_grid = 0;
    struct Slot {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        SvBatch* h_batch = nullptr;   // pinned
        SvBatch* d_batch = nullptr;
        SvResult* h_res = nullptr;    // pinned
        SvResult* d_res = nullptr;
        uint32_t* ctr = nullptr;
        hipGraphExec_t graph = nullptr;
        std::vector<uint32_t> qids;
        uint64_t block_off = 0, block_bytes = 0;   // the batch's heap block
        bool busy = false;
    };
    std::vector<Slot> slots;
};

namespace {

void sv_capture(pimdal_server* S, pimdal_server::Slot& sl) {
    pimdal_ctx* c = S->ctx;
    const uint32_t nres = 3 * SV_MAXQ;
    const uint32_t pgrid = (uint32_t)c->cu_count * 2;
    hipGraph_t g;
    PD_HIP(hipStreamBeginCapture(sl.s, hipStreamCaptureModeThreadLocal));
    PD_HIP(hipMemcpyAsync(sl.d_batch, sl.h_batch, sizeof(SvBatch), hipMemcpyHostToDevice, sl.s));
    hipLaunchKernelGGL(k_sv_reset, dim3(pgrid), dim3(256), 0, sl.s, (const SvBatch*)sl.d_batch, S->heap, sl.d_res,
                       nres, sl.ctr);
    if (S->stream_sel)
        hipLaunchKernelGGL(k_sv_stream<false>, dim3(S->stream_grid), dim3(SV2_BLOCK), 0, sl.s,
                           (const SvBatch*)sl.d_batch, S->T, S->heap, sl.d_res, sl.ctr);
    else
        hipLaunchKernelGGL(k_sv_scan<false>, dim3(pgrid), dim3(SV_BLOCK), 0, sl.s, (const SvBatch*)sl.d_batch, S->T,
                           S->heap, sl.d_res, sl.ctr);
    if (S->stream_join)
        hipLaunchKernelGGL(k_sv_stream<true>, dim3(S->stream_grid), dim3(SV2_BLOCK), 0, sl.s,
                           (const SvBatch*)sl.d_batch, S->T, S->heap, sl.d_res, sl.ctr);
    else
        hipLaunchKernelGGL(k_sv_scan<true>, dim3(pgrid), dim3(SV_BLOCK), 0, sl.s, (const SvBatch*)sl.d_batch, S->T,
                           S->heap, sl.d_res, sl.ctr);
    // the group-by's loads are 16-byte vectors with few in flight per thread:
    // eight blocks per CU keep enough of them in flight
    hipLaunchKernelGGL(k_sv_groupby, dim3(4 * pgrid), dim3(SV_BLOCK), 0, sl.s, (const SvBatch*)sl.d_batch, S->T, S->heap,
                       sl.ctr);
    hipLaunchKernelGGL(k_sv_groupby_final, dim3(SV_MAXQ), dim3(SV_G * PD_WAVE), 0, sl.s, (const SvBatch*)sl.d_batch,
                       (const uint8_t*)S->heap, (const uint8_t*)S->lut + 512, S->n1, S->n2, sl.d_res);
    PD_HIP(hipMemcpyAsync(sl.h_res, sl.d_res, sizeof(SvResult) * nres, hipMemcpyDeviceToHost, sl.s));
    PD_HIP(hipStreamEndCapture(sl.s, &g));
    PD_HIP(hipGraphInstantiate(&sl.graph, g, nullptr, nullptr, 0));
    PD_HIP(hipGraphDestroy(g));
}

double now_us(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

// Heap bytes one query takes while its batch runs: its output rows (a
// select's row ids, a join's (fact row, dim row) pairs, at most one per
// slice row), one look-back status word per 4096-row tile, and for a
// group-by one partial row per 2^16-row chunk.
struct Footprint {
    uint64_t out_bytes, status_words, items;
    uint64_t bytes() const { return out_bytes + status_words * 8 + items * SV_G * 8 * 8; }
};

Footprint footprint(const pimdal_server* S, const pimdal_query& q) {
    const uint64_t tiles = (q.rows + SV_TILE - 1) / SV_TILE;
    if (q.kind == PIMDAL_Q_SELECT) return {q.rows * 4, S->stream_sel ? sv2_tiles(q.offset, q.rows) : tiles, 0};
    if (q.kind == PIMDAL_Q_JOIN) return {q.rows * 8, S->stream_join ? sv2_tiles(q.offset, q.rows) : tiles, 0};
    return {0, 0, (q.rows + SV_GB_CHUNK - 1) / SV_GB_CHUNK};
}

// ---- the plain, one-query-at-a-time path (drop-in operator calls) ----
void run_plain(pimdal_server* S, const pimdal_query& q, pimdal_query_result& r) {
    pimdal_ctx* c = S->ctx;
    hipStream_t s = c->stream;
    const pimdal_tables& t = S->tab;
    if (q.kind == PIMDAL_Q_SELECT) {
        Scratch rows(c, std::max<uint64_t>(q.rows, 1) * 4), cnt(c, 16);
        fill_u64(c, cnt.get(), 0, 2, s);
        select_range(c, PIMDAL_I64, t.price + q.offset, nullptr, q.rows, INT64_MIN, q.param, nullptr,
                     rows.as<uint32_t>(), cnt.as<uint64_t>(), s);
        const uint64_t k = read_u64(c, cnt.get(), s);
        if (k) launch(c, "sv_ck", k_ck_rows, dim3((uint32_t)std::min<uint64_t>(ceil_div(k, 256), 1024)), dim3(256), 0, s,
                      (const uint32_t*)rows.as<uint32_t>(), k, q.offset, cnt.as<uint64_t>() + 1);
        r.count = k;
        r.checksum = read_u64(c, cnt.as<uint64_t>() + 1, s);
    } else if (q.kind == PIMDAL_Q_JOIN && S->fk_row) {
        // the resident join index of the two tables, as the batched path
        // reads it: one select over the slice of (dimension date, dimension
        // row) keeps the rows whose dimension row exists and has date < param
        // (a fact row with no dimension row carries the largest date)
        Scratch dr(c, std::max<uint64_t>(q.rows, 1) * 4), fr(c, std::max<uint64_t>(q.rows, 1) * 4), cnt(c, 16);
        fill_u64(c, cnt.get(), 0, 2, s);
        select_range(c, PIMDAL_I32, (const int32_t*)S->fk_date + q.offset, (const uint32_t*)S->fk_row + q.offset,
                     q.rows, INT64_MIN, q.param, dr.as<uint32_t>(), fr.as<uint32_t>(), cnt.as<uint64_t>(), s);
        const uint64_t k = read_u64(c, cnt.get(), s);
        if (k) launch(c, "sv_ck", k_ck_pairs, dim3((uint32_t)std::min<uint64_t>(ceil_div(k, 256), 1024)), dim3(256), 0, s,
                      (const uint32_t*)fr.as<uint32_t>(), (const uint32_t*)dr.as<uint32_t>(), k, q.offset,
                      cnt.as<uint64_t>() + 1);
        r.count = k;
        r.checksum = read_u64(c, cnt.as<uint64_t>() + 1, s);
    } else if (q.kind == PIMDAL_Q_JOIN) {
        // no resident join index: qualifying dimension rows (key, row) with
        // date < param, then the hash join with the fact slice
        Scratch dk(c, t.n_dim * 4), dr(c, t.n_dim * 4), cnt(c, 16);
        fill_u64(c, cnt.get(), 0, 2, s);
        select_range(c, PIMDAL_I32, t.dim_date, reinterpret_cast<const uint32_t*>(t.dim_key), t.n_dim, INT64_MIN,
                     q.param, dk.as<uint32_t>(), dr.as<uint32_t>(), cnt.as<uint64_t>(), s);
        const uint64_t nd = read_u64(c, cnt.get(), s);
        Scratch frow(c, std::max<uint64_t>(q.rows, 1) * 4), op(c, std::max<uint64_t>(q.rows, 1) * 4),
            ob(c, std::max<uint64_t>(q.rows, 1) * 4);
        launch(c, "sv_iota", k_iota64, dim3((uint32_t)std::min<uint64_t>(ceil_div(q.rows, 256), 4096)), dim3(256), 0, s,
               frow.as<uint32_t>(), q.rows, q.offset);
        const uint64_t k = hash_join(c, dk.as<uint32_t>(), dr.as<uint32_t>(), nd,
                                     reinterpret_cast<const uint32_t*>(t.orderkey + q.offset), frow.as<uint32_t>(),
                                     q.rows, op.as<uint32_t>(), ob.as<uint32_t>(), q.rows, s);
        if (k) launch(c, "sv_ck", k_ck_pairs, dim3((uint32_t)std::min<uint64_t>(ceil_div(k, 256), 1024)), dim3(256), 0, s,
                      (const uint32_t*)op.as<uint32_t>(), (const uint32_t*)ob.as<uint32_t>(), k, (uint64_t)0,
                      cnt.as<uint64_t>() + 1);
        r.count = k;
        r.checksum = read_u64(c, cnt.as<uint64_t>() + 1, s);
    } else {
        Scratch o(c, 64 * (4 + 8 * 6 + 16));
        int64_t* base = o.as<int64_t>();
        pimdal_q1_result out{reinterpret_cast<int32_t*>(base), base + 64, base + 128, base + 192, base + 256,
                             base + 320, base + 448};
        const uint64_t g = filter_groupby_q1(c, t.date + q.offset, t.flag1 + q.offset, t.flag2 + q.offset, nullptr,
                                             t.quantity + q.offset, t.price + q.offset, t.discount + q.offset,
                                             t.tax + q.offset, q.rows, (int32_t)std::max<int64_t>(
                                                 std::min<int64_t>(q.param, INT32_MAX), INT32_MIN),
                                             out, 64, s);
        std::vector<int64_t> h(64 * 8);
        PD_HIP(hipMemcpyAsync(h.data(), base, 64 * 8 * 8, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        const int32_t* keys = reinterpret_cast<const int32_t*>(h.data());
        uint64_t ck = 0;
        for (uint64_t i = 0; i < g; i++) {
            const int64_t v[7] = {h[64 + i], h[128 + i], h[192 + i], h[256 + i], h[320 + 2 * i], h[320 + 2 * i + 1],
                                  h[448 + i]};
            ck += pd_ck_group(keys[i], v);
        }
        r.count = g;
        r.checksum = ck;
    }
}

}  // namespace

extern "C" int pimdal_server_create(pimdal_ctx* c, const pimdal_tables* tables, int slots, uint64_t heap_bytes,
                                    pimdal_server** out) {
    PD_API_BEGIN
    PD_REQUIRE(c && tables && out, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    PD_REQUIRE(slots >= 1 && slots <= 16, PIMDAL_E_INVALID, "slots must be in [1, 16]");
    const pimdal_tables& t = *tables;
    PD_REQUIRE(t.orderkey && t.quantity && t.price && t.discount && t.tax && t.date && t.flag1 && t.flag2 && t.n_fact,
               PIMDAL_E_INVALID, "missing fact column");
    PD_REQUIRE(t.dim_key && t.dim_date && t.n_dim, PIMDAL_E_INVALID, "missing dimension column");
    PD_REQUIRE(t.n_fact < (1ull << 32) && t.n_dim < (1ull << 31), PIMDAL_E_INVALID, "tables too large for 32-bit row ids");
    auto* S = new pimdal_server();
    try {
        S->ctx = c;
        S->tab = t;
        hipStream_t s = c->stream;
        // resident dimension-key index: power-of-two slots at load <= 0.5.
        // Higher loads lengthen linear-probe chains, and a wave waits for its
        // slowest lane, so the tail of the chain length sets the probe cost.
        const double load = (double)c->tune_get("serve_index_load_pct", 50) / 100.0;
        uint64_t cap = 1;
        while ((double)cap * load < (double)t.n_dim) cap <<= 1;
        S->dim_index = c->pool.alloc(cap * 16, s);
        fill_u64(c, S->dim_index, 0, 2 * cap, s);
        Scratch dup(c, 4, s);
        fill_u32(c, dup.get(), 0, 1, s);
        launch(c, "sv_dim_index", k_dim_index_build, dim3((uint32_t)std::min<uint64_t>(ceil_div(t.n_dim, 256), 8192)),
               dim3(256), 0, s, t.dim_key, t.dim_date, t.n_dim, (uint4*)S->dim_index, cap - 1, dup.as<uint32_t>());
        uint32_t hd = 0;
        read_small(c, &hd, dup.get(), 4, s);
        PD_REQUIRE(hd == 0, PIMDAL_E_INVALID, "dimension keys must be unique");
        // flag domains of the fact table -> dense group codes (fixed while the table is resident)
        S->lut = c->pool.alloc(1024, s);
        uint32_t meta[2];
        flag_codes(c, t.flag1, t.flag2, t.n_fact, (uint8_t*)S->lut, meta, S->h_inv, s);
        S->n1 = meta[0];
        S->n2 = meta[1];
        S->T = SvTables{t.orderkey, t.quantity, t.price, t.discount, t.tax, t.date, t.flag1, t.flag2, t.n_fact,
                        t.dim_date, (const uint4*)S->dim_index, cap - 1, (const uint8_t*)S->lut, S->n2,
                        nullptr, nullptr};
        // foreign-key join index of the resident pair (fact.orderkey ->
        // dim.key): joins then stream 8 bytes per fact row instead of one
        // random index probe per row (both run modes read it)
        if (c->tune_get("serve_join_index", 1)) {
            S->fk_row = c->pool.alloc(t.n_fact * 4, s);
            S->fk_date = c->pool.alloc(t.n_fact * 4, s);
            launch(c, "sv_fk_index", k_fk_index_build, dim3((uint32_t)std::min<uint64_t>(ceil_div(t.n_fact, 256), 16384)),
                   dim3(256), 0, s, S->T, (uint32_t*)S->fk_row, (int32_t*)S->fk_date);
            S->T.fk_row = (const uint32_t*)S->fk_row;
            S->T.fk_date = (const int32_t*)S->fk_date;
        }
        // the serving heap: by default half of what the budget leaves after
        // the tables' indexes (at most 8 GiB, at least 64 MiB)
        if (!heap_bytes) {
            const uint64_t room = c->pool.budget() > c->pool.reserved() ? c->pool.budget() - c->pool.reserved() : 0;
            heap_bytes = std::max<uint64_t>(64ull << 20, std::min<uint64_t>(room / 2, 8ull << 30));
        }
        S->heap = (uint8_t*)c->pool.alloc(heap_bytes, s);
        S->heap_alloc.init(heap_bytes);
        // the streaming kernels load 16 bytes at a time from 4-row-aligned rows
        S->stream_sel = ((uintptr_t)t.price % 16) == 0;
        S->stream_join = S->fk_row != nullptr;   // pool allocations: 256-byte aligned
        S->stream_grid = blocks_per_cu(k_sv_stream<false>, SV2_BLOCK) * (uint32_t)c->cu_count;
        // slots
        S->nslots = slots;
        S->slots.resize(slots);
        for (auto& sl : S->slots) {
            PD_HIP(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
            PD_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
            PD_HIP(hipHostMalloc((void**)&sl.h_batch, sizeof(SvBatch), hipHostMallocDefault));
            PD_HIP(hipHostMalloc((void**)&sl.h_res, sizeof(SvResult) * 3 * SV_MAXQ, hipHostMallocDefault));
            memset(sl.h_batch, 0, sizeof(SvBatch));
            sl.d_batch = (SvBatch*)c->pool.alloc(sizeof(SvBatch), s);
            sl.d_res = (SvResult*)c->pool.alloc(sizeof(SvResult) * 3 * SV_MAXQ, s);
            sl.ctr = (uint32_t*)c->pool.alloc(16, s);
        }
        PD_HIP(hipStreamSynchronize(s));
        for (auto& sl : S->slots) sv_capture(S, sl);
    } catch (...) {
        pimdal_server_destroy(S);
        throw;
    }
    *out = S;
    PD_API_END
}

extern "C" int pimdal_server_destroy(pimdal_server* S) {
    PD_API_BEGIN
    if (!S) return PIMDAL_OK;
    pimdal_ctx* c = S->ctx;
    for (auto& sl : S->slots) {
        if (sl.s) (void)hipStreamSynchronize(sl.s);
        if (sl.graph) (void)hipGraphExecDestroy(sl.graph);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.h_batch) (void)hipHostFree(sl.h_batch);
        if (sl.h_res) (void)hipHostFree(sl.h_res);
        for (void* p : {(void*)sl.d_batch, (void*)sl.d_res, (void*)sl.ctr})
            if (p) c->pool.free(p, c->stream);
        if (sl.s) (void)hipStreamDestroy(sl.s);
    }
    if (S->heap) c->pool.free(S->heap, c->stream);
    if (S->dim_index) c->pool.free(S->dim_index, c->stream);
    if (S->fk_row) c->pool.free(S->fk_row, c->stream);
    if (S->fk_date) c->pool.free(S->fk_date, c->stream);
    if (S->lut) c->pool.free(S->lut, c->stream);
    delete S;
    PD_API_END
}

namespace {

// Builds one batch into slot `sl` from the tenants' queues and admits it:
// queries are taken round-robin over the tenants, FIFO within a tenant, as
// long as the batch's heap block (its queries' footprints, each region
// 256-byte aligned) fits the largest free range of the heap and the
// tenant stays within its share of the heap (heap / tenants with work,
// unless it is the only one).  A tenant whose oldest query does not fit
// stops taking queries for this batch; others go on.  Returns the queries
// admitted (0: nothing fits now).
struct TenantQ {
    std::deque<uint64_t> q;
    uint64_t inflight_bytes = 0;
};

uint64_t admit_batch(pimdal_server* S, pimdal_server::Slot& sl, const pimdal_query* qs,
                     std::map<uint32_t, TenantQ>& tenants, uint32_t& rr, std::vector<uint64_t>& qbytes,
                     pimdal_query_result* res, double now) {
    SvBatch& B = *sl.h_batch;
    B.sel.nq = B.join.nq = B.gb.nq = 0;
    B.sel.items = B.join.items = B.gb.items = 0;
    // a batch takes at most its slot's share of the heap (unless one query
    // needs more), so every slot can hold a batch and batches overlap
    const uint64_t limit = S->heap_alloc.largest();
    const uint64_t batch_cap = S->heap_alloc.size() / (uint64_t)S->nslots;
    uint32_t active = 0;
    for (auto& kv : tenants) active += (!kv.second.q.empty() || kv.second.inflight_bytes) ? 1 : 0;
    const uint64_t share = active > 1 ? S->heap_alloc.size() / active : S->heap_alloc.size();
    // sizes of the four regions as queries join the batch
    uint64_t sel_bytes = 0, join_bytes = 0, status = 0, gb_items = 0;
    auto block_bytes = [&](uint64_t sb, uint64_t jb, uint64_t st, uint64_t gi) {
        return SvHeap::round(sb) + SvHeap::round(jb) + SvHeap::round(st * 8) + SvHeap::round(gi * SV_G * 64);
    };
    std::map<uint32_t, uint64_t> taken;   // bytes per tenant in this batch
    std::vector<uint64_t> picked;
    std::vector<uint32_t> ids;
    for (auto& kv : tenants) ids.push_back(kv.first);
    std::vector<bool> blocked(ids.size(), false);
    const size_t nt = ids.size();
    size_t cur = nt ? rr % nt : 0;
    uint32_t stuck = 0;   // consecutive tenants that could not add a query
    while (nt && stuck < nt && picked.size() < (size_t)(3 * SV_MAXQ)) {
        TenantQ& T = tenants[ids[cur]];
        bool added = false;
        if (!blocked[cur] && !T.q.empty()) {
            const uint64_t qi = T.q.front();
            const pimdal_query& q = qs[qi];
            const Footprint f = footprint(S, q);
            SvKind& K = q.kind == PIMDAL_Q_SELECT ? B.sel : q.kind == PIMDAL_Q_JOIN ? B.join : B.gb;
            const uint64_t nb = block_bytes(sel_bytes + (q.kind == PIMDAL_Q_SELECT ? f.out_bytes : 0),
                                            join_bytes + (q.kind == PIMDAL_Q_JOIN ? f.out_bytes : 0),
                                            status + f.status_words, gb_items + f.items);
            const bool fits = K.nq < (uint32_t)SV_MAXQ && nb <= limit && (picked.empty() || nb <= batch_cap) &&
                              T.inflight_bytes + taken[ids[cur]] + f.bytes() <= std::max(share, f.bytes());
            if (fits) {
                T.q.pop_front();
                SvQuery& d = K.q[K.nq++];
                d.param = q.param;
                d.offset = q.offset;
                d.rows = q.rows;
                d.qid = (uint32_t)picked.size();
                d.pad = 0;
                d.item0 = K.items;
                if (q.kind == PIMDAL_Q_GROUPBY) {
                    d.out_off = 0;
                    d.status_off = 0;
                    K.items += f.items;
                    gb_items += f.items;
                } else {
                    d.status_off = status;
                    K.items += f.status_words;
                    status += f.status_words;
                    if (q.kind == PIMDAL_Q_SELECT) {
                        d.out_off = sel_bytes / 4;
                        sel_bytes += f.out_bytes;
                    } else {
                        d.out_off = join_bytes / 8;   // rebased past the select region below
                        join_bytes += f.out_bytes;
                    }
                }
                taken[ids[cur]] += f.bytes();
                qbytes[qi] = f.bytes();
                picked.push_back(qi);
                res[qi].start_us = now;
                added = true;
            } else {
                blocked[cur] = true;   // FIFO within a tenant: its later queries wait too
            }
        }
        stuck = added ? 0 : stuck + 1;
        cur = (cur + 1) % nt;
    }
    rr = nt ? (uint32_t)cur : 0;
    if (picked.empty()) return 0;
    const uint64_t bytes = block_bytes(sel_bytes, join_bytes, status, gb_items);
    uint64_t off = 0;
    PD_REQUIRE(S->heap_alloc.alloc(bytes, &off), PIMDAL_E_INTERNAL, "serving heap block vanished");
    const uint64_t join_base = SvHeap::round(sel_bytes);
    for (uint32_t i = 0; i < B.join.nq; i++) B.join.q[i].out_off += join_base / 8;
    B.arena_off = off;
    B.status_off = off + join_base + SvHeap::round(join_bytes);
    B.partials_off = B.status_off + SvHeap::round(status * 8);
    B.status_words = status;
    for (uint64_t qi : picked) tenants[qs[qi].tenant].inflight_bytes += qbytes[qi];
    sl.qids.assign(picked.begin(), picked.end());
    sl.block_off = off;
    sl.block_bytes = bytes;
    return picked.size();
}

}  // namespace

extern "C" int pimdal_server_run(pimdal_server* S, const pimdal_query* qs, uint64_t nq, int mode,
                                 pimdal_query_result* res, pimdal_serve_stats* stats) {
    PD_API_BEGIN
    PD_REQUIRE(S && (qs || !nq) && (res || !nq), PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(S->ctx);
    PD_REQUIRE(mode == PIMDAL_SERVE_SEQUENTIAL || mode == PIMDAL_SERVE_BATCHED, PIMDAL_E_INVALID, "unknown mode");
    for (uint64_t i = 0; i < nq; i++) {
        const pimdal_query& q = qs[i];
        PD_REQUIRE(q.kind >= PIMDAL_Q_SELECT && q.kind <= PIMDAL_Q_GROUPBY, PIMDAL_E_INVALID, "unknown query kind");
        PD_REQUIRE(q.rows >= 1 && q.offset + q.rows <= S->tab.n_fact, PIMDAL_E_INVALID, "query slice out of range");
        PD_REQUIRE(i == 0 || qs[i - 1].arrival_us <= q.arrival_us, PIMDAL_E_INVALID, "queries must be in arrival order");
        const Footprint f = footprint(S, q);
        PD_REQUIRE(SvHeap::round(f.out_bytes) + SvHeap::round(f.status_words * 8) + SvHeap::round(f.items * SV_G * 64) <=
                       S->heap_alloc.size(),
                   PIMDAL_E_NOMEM, "query larger than the serving heap (raise heap_bytes)");
    }
    pimdal_ctx* c = S->ctx;
    PD_HIP(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t batches = 0, max_inflight = 0;
    std::map<uint32_t, TenantQ> tenants;
    for (uint64_t i = 0; i < nq; i++) tenants[qs[i].tenant];
    S->heap_alloc.reset_peak();
    if (mode == PIMDAL_SERVE_SEQUENTIAL) {
        for (uint64_t i = 0; i < nq; i++) {
            while (now_us(t0) < qs[i].arrival_us) std::this_thread::yield();
            res[i].start_us = now_us(t0);
            run_plain(S, qs[i], res[i]);
            res[i].end_us = now_us(t0);
            res[i].latency_us = res[i].end_us - qs[i].arrival_us;
        }
        batches = nq;
        max_inflight = nq ? 1 : 0;
    } else {
        const bool dense_gb = S->n1 * S->n2 <= (uint32_t)SV_G;
        uint64_t next = 0, done = 0, inflight = 0;
        uint32_t rr = 0;
        std::vector<uint64_t> qbytes(nq, 0);
        while (done < nq) {
            bool progress = false;
            const double now = now_us(t0);
            while (next < nq && qs[next].arrival_us <= now) {
                if (qs[next].kind == PIMDAL_Q_GROUPBY && !dense_gb) {
                    // more than 8 flag groups: served by the plain operator, inline
                    const uint64_t qi = next;
                    res[qi].start_us = now_us(t0);
                    run_plain(S, qs[qi], res[qi]);
                    res[qi].end_us = now_us(t0);
                    res[qi].latency_us = res[qi].end_us - qs[qi].arrival_us;
                    done++;
                } else {
                    tenants[qs[next].tenant].q.push_back(next);
                }
                next++;
                progress = true;
            }
            // retire finished batches: results, heap block, tenant shares
            for (auto& sl : S->slots) {
                if (!sl.busy || hipEventQuery(sl.ev) != hipSuccess) continue;
                const double t_end = now_us(t0);
                for (size_t j = 0; j < sl.qids.size(); j++) {
                    const uint64_t qi = sl.qids[j];
                    pimdal_query_result& r = res[qi];
                    r.count = sl.h_res[j].count;
                    r.checksum = sl.h_res[j].checksum;
                    r.end_us = t_end;
                    r.latency_us = t_end - qs[qi].arrival_us;
                    tenants[qs[qi].tenant].inflight_bytes -= qbytes[qi];
                }
                S->heap_alloc.release(sl.block_off, sl.block_bytes);
                done += sl.qids.size();
                inflight -= sl.qids.size();
                sl.qids.clear();
                sl.busy = false;
                progress = true;
            }
            // admit arrived queries into free slots, by footprint against the heap
            for (auto& sl : S->slots) {
                if (sl.busy) continue;
                const uint64_t k = admit_batch(S, sl, qs, tenants, rr, qbytes, res, now_us(t0));
                if (!k) break;   // nothing fits (or nothing waits): wait for a batch to retire
                PD_HIP(hipGraphLaunch(sl.graph, sl.s));
                PD_HIP(hipEventRecord(sl.ev, sl.s));
                sl.busy = true;
                batches++;
                inflight += k;
                max_inflight = std::max(max_inflight, inflight);
                progress = true;
            }
            if (!progress) std::this_thread::yield();
        }
    }
    if (stats) {
        std::vector<double> lat(nq);
        double first = nq ? qs[0].arrival_us : 0, last = 0, sum = 0;
        for (uint64_t i = 0; i < nq; i++) {
            lat[i] = res[i].latency_us;
            last = std::max(last, res[i].end_us);
            sum += lat[i];
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) { return nq ? lat[std::min<uint64_t>(nq - 1, (uint64_t)(p * (nq - 1) + 0.5))] : 0.0; };
        stats->makespan_us = last - first;
        stats->qps = stats->makespan_us > 0 ? nq / (stats->makespan_us * 1e-6) : 0.0;
        stats->p50_us = pct(0.50);
        stats->p99_us = pct(0.99);
        stats->mean_us = nq ? sum / nq : 0.0;
        stats->batches = batches;
        stats->max_inflight = max_inflight;
        stats->heap_bytes = S->heap_alloc.size();
        stats->heap_peak_bytes = S->heap_alloc.peak();
        stats->tenants = tenants.size();
    }
    PD_API_END
}
