// radix_sort.hip — stable LSD radix sort of (uint32 key, payload) pairs,
// onesweep style.
//
// Reference semantics: reference/db_operators/sort.cpp:30-40 (sort of the key
// column) and pimdal/sort/sort_keyval.c:114-179 with sort_keyval_func.c:
// 255-364 (value-range partition plus per-split quicksort of (key, row id)
// pairs, unstable).  Here the order is the unique stable order by key.
//
// Kernels:
//  k_rs_hist      one read of the keys: LDS digit histograms for every pass
//                 plus the AND/OR of all keys (bits equal in every key need
//                 no pass);
//  k_rs_scan      exclusive scan of each pass's 256-bin histogram;
//  k_rs_onesweep  one pass per 8-bit digit: a block takes the next 8192-key
//                 tile from an atomic counter, ranks its keys stably with
//                 per-wave LDS match masks (ds_or of lane bits, read back,
//                 popcount below the lane), publishes its per-digit counts,
//                 resolves each digit's global offset by decoupled look-back
//                 over earlier tiles, reorders the tile in LDS by digit and
//                 writes keys and payloads out in digit runs.
#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

#include <algorithm>

namespace pd {

constexpr int RS_ITEMS = 16;                   // keys per thread; tile = block * 16
constexpr int RS_BINS = 256;
constexpr uint32_t RS_FLAG_AGG = 1u << 30;
constexpr uint32_t RS_FLAG_PRE = 2u << 30;
constexpr uint32_t RS_VALUE = (1u << 30) - 1;

struct RsPassDesc {
    uint32_t shift[4];
    uint32_t mask[4];
    uint32_t npass;
};

// Digit source: the key itself, or (HASH) the reference join hash of the
// key, so the same passes radix-partition rows by hash bits.
template <bool HASH>
__device__ __forceinline__ uint32_t rs_src(uint32_t k) {
    if constexpr (HASH) return pd::ref_hash(k);
    else return k;
}

// ---- histogram of every pass + AND/OR of the keys ----
// Lanes spread their LDS atomics over RS_HCOPY copies of the histograms
// (copy = lane mod RS_HCOPY), so lanes of one instruction that hit the same
// bin are fewer (few-bin digits would otherwise serialize on one address).
constexpr int RS_HCOPY = 8;

template <bool HASH>
__global__ __launch_bounds__(512) void k_rs_hist(const uint32_t* __restrict__ keys, uint64_t n, RsPassDesc pd,
                                                 uint32_t* __restrict__ g_hist, uint32_t* __restrict__ g_andor) {
    __shared__ uint32_t h[RS_HCOPY][4][RS_BINS];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < RS_HCOPY * 4 * RS_BINS; i += blockDim.x) (&h[0][0][0])[i] = 0;
    __syncthreads();
    auto& hc = h[pd::lane() % RS_HCOPY];
    uint32_t a = 0xFFFFFFFFu, o = 0;
    const uint64_t nvec = ((uintptr_t)keys % 16 == 0) ? n / 4 : 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint4* k4p = reinterpret_cast<const uint4*>(keys);
    auto count4 = [&](const uint4 q) {
        const uint32_t k4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            a &= k4[e];
            o |= k4[e];
            const uint32_t src = rs_src<HASH>(k4[e]);
            for (uint32_t p = 0; p < pd.npass; p++) atomicAdd(&hc[p][(src >> pd.shift[p]) & pd.mask[p]], 1u);
        }
    };
    // four 16-byte loads in flight per thread, then their atomics.  A digit
    // that all 16 keys of every active lane share (a constant high digit, a
    // narrow key range) takes one atomic for the wave instead of 16 per lane.
    uint64_t v = blockIdx.x * (uint64_t)blockDim.x + t;
    for (; v + 3 * stride < nvec; v += 4 * stride) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = k4p[v + u * stride];
        uint32_t src[16], qa = 0xFFFFFFFFu, qo = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t k4[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                a &= k4[e];
                o |= k4[e];
                src[u * 4 + e] = rs_src<HASH>(k4[e]);
                qa &= src[u * 4 + e];
                qo |= src[u * 4 + e];
            }
        }
        const uint32_t y0 = (uint32_t)__builtin_amdgcn_readfirstlane(qa);
        const uint32_t var = (qa ^ qo) | (qa ^ y0);   // bits that differ among this wave's keys
        const uint64_t act = pd::ballot(true);
        for (uint32_t p = 0; p < pd.npass; p++) {
            if (pd::ballot(((var >> pd.shift[p]) & pd.mask[p]) != 0) == 0) {
                if (pd::lane() == (uint32_t)__builtin_ctzll(act))
                    atomicAdd(&h[0][p][(y0 >> pd.shift[p]) & pd.mask[p]], 16u * pd::popc64(act));
            } else {
#pragma unroll
                for (int i = 0; i < 16; i++) atomicAdd(&hc[p][(src[i] >> pd.shift[p]) & pd.mask[p]], 1u);
            }
        }
    }
    for (; v < nvec; v += stride) count4(k4p[v]);
    for (uint64_t i = nvec * 4 + blockIdx.x * (uint64_t)blockDim.x + t; i < n; i += stride) {
        const uint32_t k = keys[i];
        a &= k;
        o |= k;
        const uint32_t src = rs_src<HASH>(k);
        for (uint32_t p = 0; p < pd.npass; p++) atomicAdd(&hc[p][(src >> pd.shift[p]) & pd.mask[p]], 1u);
    }
    __syncthreads();
    for (uint32_t i = t; i < pd.npass * RS_BINS; i += blockDim.x) {
        uint32_t v = 0;
#pragma unroll
        for (int cpy = 0; cpy < RS_HCOPY; cpy++) v += (&h[cpy][0][0])[i];
        if (v) atomicAdd(&g_hist[i], v);
    }
    // AND/OR over the block, then one atomic pair per block (device-wide
    // atomics on one word serialize, about 11 ns each)
#pragma unroll
    for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
        a &= __shfl_xor(a, d, PD_WAVE);
        o |= __shfl_xor(o, d, PD_WAVE);
    }
    __shared__ uint32_t s_ao[2][512 / PD_WAVE];
    if (pd::lane() == 0) {
        s_ao[0][t / PD_WAVE] = a;
        s_ao[1][t / PD_WAVE] = o;
    }
    __syncthreads();
    if (t == 0) {
        for (uint32_t i = 1; i < blockDim.x / PD_WAVE; i++) {
            a &= s_ao[0][i];
            o |= s_ao[1][i];
        }
        atomicAnd(&g_andor[0], a);
        atomicOr(&g_andor[1], o);
    }
}

// ---- exclusive scan of each pass's histogram (one block per pass) ----
__global__ __launch_bounds__(RS_BINS) void k_rs_scan(const uint32_t* __restrict__ g_hist, uint32_t* __restrict__ g_offs) {
    __shared__ uint32_t s_w[RS_BINS / PD_WAVE];
    const uint32_t t = threadIdx.x, p = blockIdx.x;
    const uint32_t v = g_hist[p * RS_BINS + t];
    const uint32_t inc = pd::wave_incl_scan(v);
    if (pd::lane() == PD_WAVE - 1) s_w[t / PD_WAVE] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t i = 0; i < t / PD_WAVE; i++) add += s_w[i];
    g_offs[p * RS_BINS + t] = add + inc - v;
}

// ---- reduce-then-scan pass: per-tile digit counts, then a scan over tiles ----
// The chained (onesweep) pass makes every tile wait until its predecessors
// have published; measured on one pass over 2^28 pairs, those waits cost
// 0.5 of its 1.3 ms (the same pass without them: 0.83 ms), while reading the
// keys once more to count each tile's digits up front costs 0.19 ms.  So
// the default pass is reduce-then-scan: k_rs_up counts, k_rs_tscan turns
// the counts into per-tile digit offsets, and the ranking/reorder kernel
// (k_rs_onesweep with RTS) reads its offsets instead of looking back.

// Digit counts of TILE-key tiles: cnt[tile][d] for d < 256 (rows past n are
// not counted).  AO: also the AND and OR of the tile's keys, andor_t[2 tile
// + 0/1] (bits equal in every key need no pass).
template <bool HASH, bool AO, int TILE>
__global__ __launch_bounds__(512) void k_rs_up(const uint32_t* __restrict__ kin, uint32_t n, uint32_t shift, uint32_t mask,
                                               uint32_t* __restrict__ cnt, uint32_t* __restrict__ andor_t) {
    constexpr int Q = TILE / 512;   // keys per thread
    __shared__ uint32_t h[RS_BINS];
    __shared__ uint32_t s_ao[2][512 / PD_WAVE];
    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    if (t < RS_BINS) h[t] = 0;
    // wave w reads rows [base + w*TILE/8, +TILE/8), lane-contiguous 4-byte
    // loads (any 4-byte alignment) through a range-checked buffer
    const uint32_t wave_base = blockIdx.x * (uint32_t)TILE + w * (uint32_t)(TILE / 8);
    const uint32_t wrows = n > wave_base ? min((uint32_t)(TILE / 8), n - wave_base) : 0u;
    const __amdgpu_buffer_rsrc_t rk = pd::buf_rsrc(kin + wave_base, wrows * 4u);
    uint32_t q[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) q[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + (uint32_t)i * PD_WAVE) * 4u, 0, 0);
    __syncthreads();
    uint32_t a = 0xFFFFFFFFu, o = 0;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        if (l + (uint32_t)i * PD_WAVE < wrows) {
            atomicAdd(&h[(rs_src<HASH>(q[i]) >> shift) & mask], 1u);
            if constexpr (AO) { a &= q[i]; o |= q[i]; }
        }
    }
    if constexpr (AO) {
#pragma unroll
        for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
            a &= __shfl_xor(a, d, PD_WAVE);
            o |= __shfl_xor(o, d, PD_WAVE);
        }
        if (l == 0) {
            s_ao[0][w] = a;
            s_ao[1][w] = o;
        }
    }
    __syncthreads();
    if (t < RS_BINS) cnt[(uint64_t)blockIdx.x * RS_BINS + t] = h[t];
    if (AO && t == 0) {
        for (int i = 1; i < 512 / PD_WAVE; i++) {
            a &= s_ao[0][i];
            o |= s_ao[1][i];
        }
        andor_t[2 * (uint64_t)blockIdx.x] = a;
        andor_t[2 * (uint64_t)blockIdx.x + 1] = o;
    }
}

// Scan over tiles, in place: cnt[tile][d] becomes the number of rows with
// digit d in tiles before `tile`; base[d] = rows with a digit below d.  A
// block (thread t = digit t) takes the next group of GT tiles from a
// counter, sums its digits, and resolves the group prefix by decoupled
// look-back over earlier groups (few: ntiles / GT).  The last group writes
// base.  AO: the groups also fold the tiles' AND/OR into andor[0..1].
constexpr uint32_t RS_GT = 64;
template <bool AO>
__global__ __launch_bounds__(RS_BINS) void k_rs_tscan(uint32_t* __restrict__ cnt, uint32_t ntiles,
                                                      uint32_t* __restrict__ gstatus, uint32_t* __restrict__ ctr,
                                                      uint32_t* __restrict__ base, const uint32_t* __restrict__ andor_t,
                                                      uint32_t* __restrict__ andor) {
    __shared__ uint32_t s_g, s_w[RS_BINS / PD_WAVE];
    const uint32_t t = threadIdx.x, l = pd::lane();
    if (t == 0) s_g = atomicAdd(ctr, 1u);
    __syncthreads();
    const uint32_t g = s_g;
    const uint32_t ngroups = (ntiles + RS_GT - 1) / RS_GT;
    const uint32_t t0 = g * RS_GT, t1 = min(ntiles, t0 + RS_GT);
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t i = t0; i < t1; i++) sum += cnt[(uint64_t)i * RS_BINS + t];
    pd::st_relaxed(&gstatus[(uint64_t)g * RS_BINS + t], (g == 0 ? RS_FLAG_PRE : RS_FLAG_AGG) | sum);
    uint32_t excl = 0;
    if (g > 0) {
        int64_t j = (int64_t)g - 1;
        bool done = false;
        while (!done) {
            uint32_t st[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                st[q] = (j - q >= 0) ? pd::ld_relaxed(&gstatus[(uint64_t)(j - q) * RS_BINS + t]) : RS_FLAG_PRE;
            bool stall = false;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (done || stall) continue;
                if ((st[q] >> 30) == 0) { stall = true; continue; }
                excl += st[q] & RS_VALUE;
                if ((st[q] >> 30) == 2) done = true;
                else j--;
            }
            if (stall) __builtin_amdgcn_s_sleep(1);
        }
        pd::st_relaxed(&gstatus[(uint64_t)g * RS_BINS + t], RS_FLAG_PRE | (excl + sum));
    }
    uint32_t run = excl;
    for (uint32_t i = t0; i < t1; i++) {
        const uint64_t k = (uint64_t)i * RS_BINS + t;
        const uint32_t c = cnt[k];
        cnt[k] = run;
        run += c;
    }
    if (g == ngroups - 1) {
        // run = total of digit t: exclusive scan over digits
        const uint32_t inc = pd::wave_incl_scan(run);
        if (l == PD_WAVE - 1) s_w[t / PD_WAVE] = inc;
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t i = 0; i < t / PD_WAVE; i++) add += s_w[i];
        base[t] = add + inc - run;
    }
    if constexpr (AO) {
        // AND/OR of the group's tiles, one atomic pair per group
        uint32_t a = 0xFFFFFFFFu, o = 0;
        for (uint32_t i = t0 + t; i < t1; i += RS_BINS) {
            a &= andor_t[2 * (uint64_t)i];
            o |= andor_t[2 * (uint64_t)i + 1];
        }
#pragma unroll
        for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
            a &= __shfl_xor(a, d, PD_WAVE);
            o |= __shfl_xor(o, d, PD_WAVE);
        }
        if (l == 0 && (t / PD_WAVE) * PD_WAVE < t1 - t0) {
            atomicAnd(&andor[0], a);
            atomicOr(&andor[1], o);
        }
    }
}

// ---- one onesweep pass ----
// Compiler-only barrier: LDS instructions of one wave execute in issue order,
// so the ranking steps below need ordering against code motion, not waits.
// (Plain LDS accesses between these barriers, not volatile ones: the
// compiler follows a volatile access with a wait for every outstanding
// global load and store too, which would stall the ranking on the payload
// loads.)
#define PD_COMPILER_BARRIER() __asm__ volatile("" ::: "memory")

// 4 waves per SIMD: at most 128 VGPRs (second launch-bounds argument is the
// minimum number of waves per execution unit).  EARLY: resolve the global
// offsets (look-back) before the LDS reorder instead of after it.
// BALLOT: rank with per-bit ballots instead of LDS match masks.
// LBW: statuses each digit's thread reads per look-back round trip.  A
// round trip costs about a microsecond and a tile walks back over a few
// dozen published aggregates, yet wide windows are slower: the cost of a
// round grows with the statuses it reads, and the status reads of all tiles
// in flight compete with the pass's own traffic.
// RTS (reduce-then-scan): tile = blockIdx.x, no status words and no
// look-back: `offs` is the scanned per-tile counts (k_rs_tscan) and `base`
// the digit starts, so no tile waits on another.
// SPLIT: every thread of the block takes part in the look-back, BLOCK / 256
// threads per digit, each reading LBW consecutive statuses of one window, so
// one round trip covers BLOCK / 256 * LBW tiles; the parts meet in LDS after
// a barrier.
// NT: bit 0 = non-temporal input loads, bit 1 = non-temporal output stores
// (the pass touches every key and payload once).
// XP (XCD-persistent): a co-resident grid (cooperative launch) whose blocks
// loop over tiles.  Block b is in class x = b % 8 (workgroups go to the 8
// XCDs round-robin, so a class is one XCD's blocks); block i of the class's
// B blocks takes the class's tickets i, i + B, i + 2B, ..., and ticket c is
// tile (c / S) * 8S + x * S + c % S with S = xcd_tiles.  Each XCD thus
// processes runs of S consecutive tiles, and the partial 64-byte lines where
// neighbouring tiles' digit runs meet are merged in one L2 before they are
// written back (runs of 16 tiles: 1.06 ms per 2^28-pair pass against 1.26
// ms with tiles dealt to blocks in launch order).  Every tile waits only on
// lower tiles.  The lowest unfinished tile belongs to one block, whose
// earlier tickets are lower tiles and so finished; with all blocks resident
// the pass always progresses.
template <typename V, bool HAS_VALS, int BLOCK, bool EARLY, bool HASH, bool BALLOT = false, int LBW = 4,
          int ITEMS = RS_ITEMS, bool RTS = false, bool SPLIT = false, int NT = 0, bool XP = false>
__global__ __launch_bounds__(BLOCK, ITEMS <= 8 ? 8 : 4) void k_rs_onesweep(
    const uint32_t* __restrict__ kin, const V* __restrict__ vin, uint32_t* __restrict__ kout, V* __restrict__ vout,
    uint64_t n, uint32_t shift, uint32_t mask, const uint32_t* __restrict__ offs, uint32_t* __restrict__ status,
    uint32_t* __restrict__ tile_ctr, const uint32_t* __restrict__ base, uint32_t xcd_tiles,
    uint32_t* __restrict__ status_next) {
    constexpr int WAVES = BLOCK / PD_WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(TILE * 4 >= WAVES * RS_BINS * 8, "match masks must fit in the key buffer");
    static_assert(!(RTS && (EARLY || BALLOT)), "the counted pass has no look-back");
    static_assert(!SPLIT || (!RTS && !EARLY && BLOCK % RS_BINS == 0), "split look-back: after the reorder");
    constexpr int PARTS = SPLIT ? BLOCK / RS_BINS : 1;
    __shared__ uint32_t s_tile;
    // keys reordered by digit; the per-wave match masks live here during ranking
    __shared__ __align__(16) uint32_t s_keys[TILE];
    __shared__ __align__(16) V s_vals[HAS_VALS ? TILE : 1];
    __shared__ uint16_t s_cnt[WAVES][RS_BINS];
    __shared__ uint32_t s_off[RS_BINS];   // tile-local digit start
    __shared__ uint32_t s_gbase[RS_BINS]; // global position of the digit's run minus its local start
    __shared__ uint32_t s_wsum[RS_BINS / PD_WAVE];
    // split look-back exchange, double-buffered by round: part sums, part
    // tiles consumed | kind << 6, and per wave whether any digit still walks
    __shared__ uint32_t s_lsum[SPLIT ? 2 : 1][PARTS][SPLIT ? RS_BINS : 1];
    __shared__ uint8_t s_lpc[SPLIT ? 2 : 1][PARTS][SPLIT ? RS_BINS : 1];
    __shared__ uint32_t s_busy[2][WAVES];
    auto s_match = reinterpret_cast<uint64_t (*)[RS_BINS]>(s_keys);

    static_assert(!XP || (!RTS && !EARLY && !SPLIT), "XCD-persistent passes are chained, look-back after the reorder");
    // XP: `offs` is this pass's 256-bin digit histogram; each block scans it
    // once into s_offs (the digit starts) instead of a separate scan kernel
    __shared__ uint32_t s_offs[XP ? RS_BINS : 1];
    if constexpr (XP) {
        const uint32_t t = threadIdx.x, w = t / PD_WAVE;
        uint32_t h = 0, inc = 0;
        if (t < RS_BINS) {
            h = offs[t];
            inc = pd::wave_incl_scan(h);
            if (pd::lane() == PD_WAVE - 1) s_wsum[w] = inc;
        }
        __syncthreads();
        if (t < RS_BINS) {
            uint32_t add = 0;
            for (uint32_t i = 0; i < w; i++) add += s_wsum[i];
            s_offs[t] = add + inc - h;
        }
        // (the first tile's barriers order these writes before any read)
    }
    // XP: block i of class x's B blocks takes the class's tickets i, i + B,
    // i + 2B, ... (no atomics); ticket c -> tile (c / S) * 8S + x * S + c % S
    const uint32_t xp_x = blockIdx.x % 8, xp_s = xcd_tiles;
    auto xp_tile = [&](uint32_t cc) -> uint32_t { return (cc / xp_s) * (8 * xp_s) + xp_x * xp_s + cc % xp_s; };
    uint32_t xp_iter = 0;
    for (;;) {   // one tile per iteration; one iteration unless XP
    // thread and lane ids through opaque copies: values derived from them
    // are recomputed per tile instead of being held across the loop
    uint32_t t = threadIdx.x, l = pd::lane();
    if constexpr (XP) {
        asm volatile("" : "+v"(t));
        asm volatile("" : "+v"(l));
    }
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    if (t == 0) {
        if constexpr (XP)
            s_tile = xp_tile(blockIdx.x / 8 + xp_iter * (gridDim.x / 8));
        else if (!RTS || tile_ctr)
            s_tile = atomicAdd(tile_ctr, 1u);
    }
    for (uint32_t i = l; i < RS_BINS; i += PD_WAVE) {
        s_match[w][i] = 0;
        s_cnt[w][i] = 0;
    }
    __syncthreads();
    // RTS tile order: ticket (tile_ctr), XCD-contiguous (xcd_tiles > 0:
    // workgroups go to the 8 XCDs round-robin, so block b on XCD b % 8 takes
    // tile (b % 8) * xcd_tiles + b / 8 and neighbouring tiles share an L2),
    // or block order
    const uint64_t tile = (!RTS || tile_ctr) ? (uint64_t)s_tile
                          : xcd_tiles ? (uint64_t)(blockIdx.x % 8) * xcd_tiles + blockIdx.x / 8
                                      : (uint64_t)blockIdx.x;
    if (RTS && tile * TILE >= n) return;   // XCD order: grid rounded up to a multiple of 8
    if (XP && tile * TILE >= n) break;     // this counter's tiles are done (s_tile is block-wide)
    // the next pass's status words of this tile start at zero
    if (status_next && t < RS_BINS) status_next[tile * RS_BINS + t] = 0;
    // RTS: this tile's run start of digit t, loaded now (used after the reorder)
    uint32_t rts_start = 0;
    if (RTS && t < RS_BINS) rts_start = base[t] + offs[tile * RS_BINS + t];
    const uint64_t tile_base = tile * TILE;
    const uint64_t wave_base = tile_base + (uint64_t)w * (PD_WAVE * ITEMS);
    const uint32_t tile_count = (uint32_t)min((uint64_t)TILE, n - tile_base);

    // ---- load keys and payloads, wave-striped: idx = wave_base + i*64 + lane ----
    // Rows past n get the largest digit, so they rank after every real row
    // of the last tile and land past n (never written).
    uint32_t key[ITEMS];
    V val[HAS_VALS ? ITEMS : 1];
    // rows of this wave that exist (wave-uniform), checked with 32-bit offsets
    const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(
        (uint32_t)(n > wave_base ? min((uint64_t)(PD_WAVE * ITEMS), n - wave_base) : 0));
    // buffer loads whose range ends at the wave's last row (0 past it)
    const __amdgpu_buffer_rsrc_t rk = pd::buf_rsrc(kin + wave_base, wrows * 4u);
#pragma unroll
    for (int i = 0; i < ITEMS; i++)
        key[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + (uint32_t)i * PD_WAVE) * 4u, 0, (NT & 1) ? 2 : 0);
    if constexpr (HAS_VALS) {
        const __amdgpu_buffer_rsrc_t rv = pd::buf_rsrc(vin + wave_base, wrows * (uint32_t)sizeof(V));
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            if constexpr (sizeof(V) == 4)
                val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, (l + (uint32_t)i * PD_WAVE) * 4u, 0, (NT & 1) ? 2 : 0);
            else
                val[i] = __builtin_bit_cast(
                    V, __builtin_amdgcn_raw_buffer_load_b64(rv, (l + (uint32_t)i * PD_WAVE) * 8u, 0, (NT & 1) ? 2 : 0));
        }
    }
    // digit of item i, computed where it is used: the key goes through an
    // opaque copy, so the compiler cannot compute all digits up front (and
    // hold them in registers beside the keys)
    auto digit_of = [&](int i) -> uint32_t {
        uint32_t kk = key[i];
        asm volatile("" : "+v"(kk));
        return ((uint32_t)i * PD_WAVE + l) < wrows ? ((rs_src<HASH>(kk) >> shift) & mask) : (RS_BINS - 1);
    };

    // ---- stable rank inside the wave ----
    const uint64_t lt = pd::lanemask_lt();
    uint32_t pos[ITEMS];
    if constexpr (BALLOT) {
        // lanes holding the same digit = AND over the digit's bits of the
        // bit's ballot (or its complement); the lowest such lane bumps the
        // wave's running count of the digit
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t d = digit_of(i);
            uint64_t m = ~0ull;
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint64_t bb = pd::ballot((d >> b) & 1);
                m &= ((d >> b) & 1) ? bb : ~bb;
            }
            const uint32_t below = pd::popc64(m & lt);
            const uint32_t old = s_cnt[w][d];
            pos[i] = (old + below) | (d << 16);
            PD_COMPILER_BARRIER();
            if (below == 0) s_cnt[w][d] = (uint16_t)(old + pd::popc64(m));
            PD_COMPILER_BARRIER();
        }
    } else
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t d = digit_of(i);
        atomicOr((unsigned long long*)&s_match[w][d], (unsigned long long)(1ull << l));
        PD_COMPILER_BARRIER();
        const uint64_t m = s_match[w][d];
        const uint32_t old = s_cnt[w][d];
        const uint32_t below = pd::popc64(m & lt);
        pos[i] = (old + below) | (d << 16);
        PD_COMPILER_BARRIER();
        if (below == 0) {
            s_cnt[w][d] = (uint16_t)(old + pd::popc64(m));
            s_match[w][d] = 0;
        }
        PD_COMPILER_BARRIER();
    }
    __syncthreads();

    // per-digit look-back (thread t = digit t): exclusive prefix of this
    // digit over all earlier tiles, LBW statuses per round trip
    auto lookback = [&](uint32_t cnt) -> uint32_t {
        uint32_t excl = 0;
        int64_t j = (int64_t)tile - 1;
        bool done = false;
        while (!done) {
            uint32_t st[LBW];
#pragma unroll
            for (int q = 0; q < LBW; q++)
                st[q] = (j - q >= 0) ? pd::ld_relaxed(&status[(uint64_t)(j - q) * RS_BINS + t]) : RS_FLAG_PRE;
            bool stall = false;
#pragma unroll
            for (int q = 0; q < LBW; q++) {
                if (done || stall) continue;
                if ((st[q] >> 30) == 0) { stall = true; continue; }
                excl += st[q] & RS_VALUE;
                if ((st[q] >> 30) == 2) done = true;
                else j--;
            }
            if (stall) __builtin_amdgcn_s_sleep(1);
        }
        pd::st_relaxed(&status[tile * RS_BINS + t], RS_FLAG_PRE | (excl + cnt));
        return excl;
    };

    // split look-back, called by every thread: thread t reads statuses
    // j - (t / 256) * LBW - q of digit t % 256.  A part ends at its first
    // inclusive prefix (counted, kind 1) or first unpublished status (not
    // counted, kind 2); the parts are combined in window order.  Every digit's
    // threads take the same steps, and the loop ends for the whole block
    // once no wave has a digit still walking.
    auto lookback_split = [&](uint32_t cnt) -> uint32_t {
        const uint32_t d = t % RS_BINS, h = t / RS_BINS;
        uint32_t excl = 0;
        int64_t j = (int64_t)tile - 1;
        bool done = tile == 0;
        for (uint32_t r = 0;; r++) {
            const uint32_t b = r & 1;
            uint32_t psum = 0, pc = 0;
            if (!done) {
                const int64_t j0 = j - (int64_t)h * LBW;
                uint32_t st[LBW];
#pragma unroll
                for (int q = 0; q < LBW; q++)
                    st[q] = (j0 - q >= 0) ? pd::ld_relaxed(&status[(uint64_t)(j0 - q) * RS_BINS + d]) : RS_FLAG_PRE;
                uint32_t kind = 0, used = 0;
#pragma unroll
                for (int q = 0; q < LBW; q++) {
                    if (kind) continue;
                    const uint32_t f = st[q] >> 30;
                    if (f == 0) { kind = 2; continue; }
                    psum += st[q] & RS_VALUE;
                    used++;
                    if (f == 2) kind = 1;
                }
                pc = used | (kind << 6);
            }
            s_lsum[b][h][d] = psum;
            s_lpc[b][h][d] = (uint8_t)pc;
            const uint64_t busy = pd::ballot(!done);
            if (l == 0) s_busy[b][w] = busy != 0;
            __syncthreads();
            uint32_t any = 0;
#pragma unroll
            for (int ww = 0; ww < WAVES; ww++) any |= s_busy[b][ww];
            if (!any) break;
            if (!done) {
                uint32_t sum = 0, used = 0, kind = 0;
#pragma unroll
                for (int p = 0; p < PARTS; p++) {
                    if (kind) break;
                    const uint32_t pcp = s_lpc[b][p][d];
                    sum += s_lsum[b][p][d];
                    used += pcp & 63u;
                    kind = pcp >> 6;
                }
                excl += sum;
                j -= used;
                if (kind == 1) done = true;
                else if (kind == 2) __builtin_amdgcn_s_sleep(1);
            }
        }
        if (h == 0 && tile > 0) pd::st_relaxed(&status[tile * RS_BINS + d], RS_FLAG_PRE | (excl + cnt));
        return excl;
    };

    // ---- per digit: prefix over waves, tile count, publish the aggregate ----
    uint32_t cnt_d = 0, inc_d = 0, excl = 0;
    if (t < RS_BINS) {
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) {
            const uint32_t c = s_cnt[ww][t];
            s_cnt[ww][t] = (uint16_t)cnt_d;
            cnt_d += c;
        }
        if constexpr (!RTS) {
            pd::st_relaxed(&status[tile * RS_BINS + t], (tile == 0 ? RS_FLAG_PRE : RS_FLAG_AGG) | cnt_d);
            if (EARLY && tile > 0) excl = lookback(cnt_d);
        }
        inc_d = pd::wave_incl_scan(cnt_d);
        if (l == PD_WAVE - 1) s_wsum[w] = inc_d;
    }
    __syncthreads();
    uint32_t local_d = 0;
    if (t < RS_BINS) {
        uint32_t add = 0;
        for (uint32_t i = 0; i < w; i++) add += s_wsum[i];
        local_d = add + inc_d - cnt_d;
        s_off[t] = local_d;
        if (EARLY) s_gbase[t] = offs[t] + excl - local_d;
    }
    __syncthreads();

    // ---- reorder the tile by digit in LDS (keys and payloads) ----
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const uint32_t d = pos[i] >> 16;
        const uint32_t p = (pos[i] & 0xFFFFu) + s_off[d] + s_cnt[w][d];
        s_keys[p] = key[i];
        if constexpr (HAS_VALS) s_vals[p] = val[i];
    }
    if (RTS && t < RS_BINS) s_gbase[t] = rts_start - local_d;
    if constexpr (SPLIT) {
        excl = lookback_split(cnt_d);
        if (t < RS_BINS) s_gbase[t] = offs[t] + excl - local_d;
    } else if (!RTS && !EARLY && t < RS_BINS) {
        if (tile > 0) excl = lookback(cnt_d);
        s_gbase[t] = (XP ? s_offs[t] : offs[t]) + excl - local_d;
    }
    __syncthreads();

    // ---- write keys and payloads in digit runs ----
#pragma unroll 4
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t j = t + k * BLOCK;
        if (j >= tile_count) break;   // slots past tile_count hold padding rows
        const uint32_t kk = s_keys[j];
        const uint64_t dst = (uint64_t)s_gbase[(rs_src<HASH>(kk) >> shift) & mask] + j;
        if (dst < n) {
            if constexpr ((NT & 2) != 0) {
                __builtin_nontemporal_store(kk, kout + dst);
                if constexpr (HAS_VALS) __builtin_nontemporal_store(s_vals[j], vout + dst);
            } else {
                kout[dst] = kk;
                if constexpr (HAS_VALS) vout[dst] = s_vals[j];
            }
        }
    }
    if constexpr (!XP) break;
    xp_iter++;
    __syncthreads();   // the next tile reuses the LDS buffers and s_tile
    }

}

// Look-back window of the passes (knob sort_lbw: 1, 2, 4, 8, 16 or 32;
// default 2: wider windows read statuses the walk does not need, and the
// status reads of ~500 tiles in flight compete with the pass's own traffic).
static int sort_lbw(pimdal_ctx* c) {
    const int64_t v = c->tune_get("sort_lbw", 2);
    return v >= 32 ? 32 : v >= 16 ? 16 : v >= 8 ? 8 : v >= 4 ? 4 : v >= 2 ? 2 : 1;
}

template <typename V, bool HV, bool HASH = false>
static void run_passes(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                       const RsPassDesc& pdsc, const uint32_t* hist, uint32_t* offs, uint32_t* ctrs, hipStream_t s) {
    const int64_t bk = c->tune_get("sort_block", 512);
    const int block = bk >= 1024 ? 1024 : bk <= 256 ? 256 : 512;
    const bool early = c->tune_get("sort_early_lookback", 0) != 0;
    // keys per thread: 16, or 8 (1024-thread blocks always take 8, so a
    // tile stays 8192 keys and two blocks still fit a CU's LDS)
    const int items = (block == 1024 || c->tune_get("sort_items", 16) == 8) ? 8 : 16;
    const bool ballot = c->tune_get("sort_rank", 0) == 1;
    const int lbw = sort_lbw(c);
    // knob sort_lb_split: 1 = two threads per digit in the look-back
    const bool split = c->tune_get("sort_lb_split", 0) == 1;
    // knob sort_nt: non-temporal loads (1), stores (2) or both (3)
    const int nt = (int)(c->tune_get("sort_nt", 0) & 3);
    const uint64_t ntiles = ceil_div(n, (uint64_t)block * items);
    const uint32_t P = pdsc.npass;
    Scratch tk(c, P > 1 ? n * 4 : 0, s);
    Scratch tv(c, (HV && P > 1) ? n * sizeof(V) : 0, s);
    // knob sort_xcd_tiles S > 0 (default 16): XCD-persistent passes with
    // runs of S tiles per XCD (512 x 16 and 256 x 16 shapes, look-back
    // after the reorder); 0 = one block per tile in launch order
    const int64_t xs_knob = std::min<int64_t>(c->tune_get("sort_xcd_tiles", 16), 4096);
    const bool xp_shape = !early && !split && nt == 0 &&
                          ((block == 512 && items == 16) || (block == 256 && items == 16 && !ballot));
    // cleared if the runtime refuses the co-resident grid; never while a plan
    // is captured (cooperative launches are not recorded into plans), nor
    // for inputs of fewer tiles than the grid has blocks (a small sort
    // would pay for launching the whole resident grid every pass)
    bool xp = xp_shape && xs_knob > 0 && !c->capture && ntiles >= (uint64_t)c->cu_count * 2;
    auto xs_of = [&](uint32_t) -> uint32_t { return xp ? (uint32_t)xs_knob : 0u; };
    // status words: XCD-persistent passes scan their digit histogram
    // themselves and clear the next pass's statuses (two buffers in turn);
    // the others take the scanned offsets and a cleared buffer per pass
    const uint64_t nst = ntiles * RS_BINS;
    Scratch st(c, (xp && P > 1 ? 2 : 1) * nst * 4, s);
    if (xp) fill_u32(c, st.get(), 0, nst, s);
    else launch(c, "radix_scan", k_rs_scan, dim3(P), dim3(RS_BINS), 0, s, hist, offs);
    // co-resident grid of an XCD-persistent pass: a multiple of 8 blocks
    auto xp_grid_of = [&](auto kernel) -> uint32_t {
        return std::max<uint32_t>(8, (blocks_per_cu(kernel, block) * c->cu_count) / 8 * 8);
    };
    const uint32_t* ksrc = kin;
    const V* vsrc = vin;
    const char* name = HASH ? "part_onesweep" : "radix_onesweep";
    for (uint32_t j = 0; j < P; j++) {
        const bool to_out = ((P - 1 - j) % 2) == 0;
        uint32_t* kdst = to_out ? kout : tk.as<uint32_t>();
        V* vdst = to_out ? vout : tv.as<V>();
        if (!xp) fill_u32(c, st.get(), 0, nst, s);
        uint32_t* st_j = st.as<uint32_t>() + (xp ? (uint64_t)(j & 1) * nst : 0);
        uint32_t* st_next = (xp && j + 1 < P) ? st.as<uint32_t>() + (uint64_t)((j + 1) & 1) * nst : nullptr;
        bool refused = false;   // the co-resident grid was refused: this and later passes take one block per tile
#define PD_RS_LAUNCH(...)                                                                                         \
    launch(c, name, k_rs_onesweep<__VA_ARGS__>, dim3((uint32_t)ntiles), dim3(block), 0, s, ksrc, vsrc, kdst, vdst, n, \
           pdsc.shift[j], pdsc.mask[j], offs + j * RS_BINS, st.as<uint32_t>(), ctrs + j, (const uint32_t*)nullptr, 0u, \
           (uint32_t*)nullptr)
        if (xs_of(j) > 0)
        {
#define PD_RS_XP(B, IT, W, BAL)                                                                                   \
    refused = !launch_coop(c, name, k_rs_onesweep<V, HV, B, false, HASH, BAL, W, IT, false, false, 0, true>,         \
                dim3(xp_grid_of(k_rs_onesweep<V, HV, B, false, HASH, BAL, W, IT, false, false, 0, true>)), dim3(B),      \
                0, s, ksrc, vsrc, kdst, vdst, n, pdsc.shift[j], pdsc.mask[j], hist + j * RS_BINS, st_j,                \
                (uint32_t*)nullptr, (const uint32_t*)nullptr, xs_of(j), st_next)
            // knob sort_xcd_refuse_pass j (tests): take pass j's launch as refused
            if ((int64_t)j == c->tune_get("sort_xcd_refuse_pass", -1)) refused = true;
            else if (block == 256) PD_RS_XP(256, 16, 2, false);
            else if (ballot) PD_RS_XP(512, 16, 2, true);
            else if (lbw == 2) PD_RS_XP(512, 16, 2, false);
            else if (lbw == 1) PD_RS_XP(512, 16, 1, false);
            else if (lbw == 8) PD_RS_XP(512, 16, 8, false);
            else PD_RS_XP(512, 16, 4, false);
#undef PD_RS_XP
            if (refused) {
                // scanned offsets and a cleared status buffer, as the
                // one-block-per-tile passes take them
                xp = false;
                launch(c, "radix_scan", k_rs_scan, dim3(P), dim3(RS_BINS), 0, s, hist, offs);
                fill_u32(c, st.get(), 0, nst, s);
            }
        }
        if (xs_of(j) > 0) {
            // the XCD-persistent pass ran above
        } else if (block == 1024) PD_RS_LAUNCH(V, HV, 1024, false, HASH, false, 4, 8);
        else if (items == 8 && block == 512) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 8);
        else if (block == 512 && !early && !ballot && split) {
            if (lbw >= 8) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 8, 16, false, true);
            else PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 16, false, true);
        } else if (block == 512 && !early && !ballot && nt) {
            if (nt == 1) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 16, false, false, 1);
            else if (nt == 2) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 16, false, false, 2);
            else PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 16, false, false, 3);
        } else if (block == 512 && !early && !ballot) {
            if (lbw == 2) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 2);
            else if (lbw == 32) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 32);
            else if (lbw == 16) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 16);
            else if (lbw == 8) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 8);
            else PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4);
        } else if (block == 512 && ballot && !early) PD_RS_LAUNCH(V, HV, 512, false, HASH, true, 16);
        else if (block == 512) PD_RS_LAUNCH(V, HV, 512, true, HASH, false, 16);
        else if (early) PD_RS_LAUNCH(V, HV, 256, true, HASH, false, 16);
        else PD_RS_LAUNCH(V, HV, 256, false, HASH, false, 16);
#undef PD_RS_LAUNCH
        ksrc = kdst;
        vsrc = vdst;
    }
}

// ---- reduce-then-scan passes ----
// Scratch of the counting passes: per-tile counts (scanned in place into
// tile offsets), group statuses of the tile scan, group tickets (one per
// count), digit starts, and per-tile AND/OR words with their fold.
struct RtsBufs {
    Scratch cnt, gst, small, ao_t;
    uint32_t ntiles, ngroups;
    uint32_t* ctrs;    // 8 tickets
    uint32_t* base;    // 256
    uint32_t* andor;   // 2
    RtsBufs(pimdal_ctx* c, uint64_t n, bool ao, hipStream_t s)
        : cnt(c, ceil_div(n, 8192) * RS_BINS * 4, s), gst(c, ceil_div(ceil_div(n, 8192), RS_GT) * RS_BINS * 4, s),
          small(c, (8 + RS_BINS + 2) * 4, s), ao_t(c, ao ? ceil_div(n, 8192) * 8 : 0, s) {
        ntiles = (uint32_t)ceil_div(n, 8192);
        ngroups = (uint32_t)ceil_div(ntiles, RS_GT);
        ctrs = small.as<uint32_t>();
        base = ctrs + 8;
        andor = base + RS_BINS;
        fill_u32(c, ctrs, 0, 8, s);
        fill_u32(c, andor, 0xFFFFFFFFu, 1, s);
        fill_u32(c, andor + 1, 0, 1, s);
    }
};

// Counts one digit of every tile and scans the counts (ticket k of the
// tile scan); AO: also the AND/OR of all keys into B.andor.
template <bool HASH, bool AO>
static void rts_count(pimdal_ctx* c, const uint32_t* keys, uint64_t n, uint32_t shift, uint32_t mask, RtsBufs& B,
                      int k, hipStream_t s) {
    launch(c, HASH ? "part_count" : "radix_count", k_rs_up<HASH, AO, 8192>, dim3(B.ntiles), dim3(512), 0, s, keys,
           (uint32_t)n, shift, mask, B.cnt.as<uint32_t>(), AO ? B.ao_t.as<uint32_t>() : (uint32_t*)nullptr);
    fill_u32(c, B.gst.get(), 0, (uint64_t)B.ngroups * RS_BINS, s);
    launch(c, HASH ? "part_tscan" : "radix_tscan", k_rs_tscan<AO>, dim3(B.ngroups), dim3(RS_BINS), 0, s,
           B.cnt.as<uint32_t>(), B.ntiles, B.gst.as<uint32_t>(), B.ctrs + k, B.base,
           (const uint32_t*)(AO ? B.ao_t.as<uint32_t>() : nullptr), AO ? B.andor : (uint32_t*)nullptr);
}

// Passes over the digits of pdsc: count + scan (unless the first digit is
// already counted in B), then rank and reorder each 8192-row tile to its
// offsets.  The shape knobs pick 512 threads x 16 keys (default) or 1024 x 8.
template <typename V, bool HV, bool HASH = false>
static void run_passes_rts(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                           const RsPassDesc& pdsc, RtsBufs& B, bool first_counted, hipStream_t s) {
    const bool b1024 = c->tune_get("sort_block", 512) >= 1024;
    const uint32_t P = pdsc.npass;
    Scratch tk(c, P > 1 ? n * 4 : 0, s);
    Scratch tv(c, (HV && P > 1) ? n * sizeof(V) : 0, s);
    const uint32_t* ksrc = kin;
    const V* vsrc = vin;
    const char* name = HASH ? "part_onesweep" : "radix_onesweep";
    for (uint32_t j = 0; j < P; j++) {
        const bool to_out = ((P - 1 - j) % 2) == 0;
        uint32_t* kdst = to_out ? kout : tk.as<uint32_t>();
        V* vdst = to_out ? vout : tv.as<V>();
        if (j > 0 || !first_counted) rts_count<HASH, false>(c, ksrc, n, pdsc.shift[j], pdsc.mask[j], B, 1 + (int)j, s);
        const uint32_t* off = B.cnt.as<uint32_t>();
        // tile order (knob sort_rts_order): 0 block order, 1 ticket, 2 XCD-contiguous
        const int64_t order = c->tune_get("sort_rts_order", 2);
        uint32_t* tctr = order == 1 ? B.ctrs + 5 + (j % 3) : nullptr;
        if (order == 1) fill_u32(c, tctr, 0, 1, s);
        const uint32_t xt = order == 2 ? (uint32_t)ceil_div(B.ntiles, 8) : 0u;
        const uint32_t grid = order == 2 ? 8 * xt : B.ntiles;
        if (b1024)
            launch(c, name, k_rs_onesweep<V, HV, 1024, false, HASH, false, 4, 8, true>, dim3(grid), dim3(1024), 0, s,
                   ksrc, vsrc, kdst, vdst, n, pdsc.shift[j], pdsc.mask[j], off, (uint32_t*)nullptr, tctr,
                   (const uint32_t*)B.base, xt, (uint32_t*)nullptr);
        else
            launch(c, name, k_rs_onesweep<V, HV, 512, false, HASH, false, 4, 16, true>, dim3(grid), dim3(512), 0, s,
                   ksrc, vsrc, kdst, vdst, n, pdsc.shift[j], pdsc.mask[j], off, (uint32_t*)nullptr, tctr,
                   (const uint32_t*)B.base, xt, (uint32_t*)nullptr);
        ksrc = kdst;
        vsrc = vdst;
    }
}

// Pass scheme (knob sort_scheme): 0 = chained onesweep per range (default),
// 1 = reduce-then-scan.  Other tile shapes than 8192 keys run chained.
static bool use_rts(pimdal_ctx* c) {
    if (c->tune_get("sort_scheme", 0) != 1) return false;
    const int64_t bk = c->tune_get("sort_block", 512);
    if (bk <= 256) return false;
    if (bk < 1024 && c->tune_get("sort_items", 16) == 8) return false;
    return c->tune_get("sort_early_lookback", 0) == 0 && c->tune_get("sort_rank", 0) == 0;
}

template <typename V, bool HV>
static void sort_rts(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                     bool detect, const RsPassDesc& all, hipStream_t s) {
    RtsBufs B(c, n, detect, s);
    RsPassDesc run = all;
    bool counted = false;
    if (detect) {
        // count digit 0 and fold the AND/OR of all keys (a bit equal in
        // every key needs no pass)
        rts_count<false, true>(c, kin, n, 0, 0xFF, B, 0, s);
        uint32_t ao[2];
        read_small(c, ao, B.andor, 8, s);
        const uint32_t varying = ao[0] ^ ao[1];
        run.npass = 0;
        for (int p = 0; p < 4; p++) {
            if ((varying >> (8 * p)) & 0xFF) {
                run.shift[run.npass] = 8 * p;
                run.mask[run.npass] = 0xFF;
                run.npass++;
            }
        }
        counted = run.npass > 0 && run.shift[0] == 0;
    }
    if (run.npass == 0) {
        copy_d2d(c, kout, kin, n * 4, s);
        if (HV) copy_d2d(c, vout, vin, n * sizeof(V), s);
        return;
    }
    run_passes_rts<V, HV>(c, kin, vin, n, kout, vout, run, B, counted, s);
}

// ---- bucket pass: the last 1-2 low digits sorted in LDS per bucket ----
// After stable passes over the two highest varying digits (a, then b), rows
// with equal high bits (bucket = digit a | digit b << 8, 2^16 buckets in key
// order) are contiguous and keep their input order.  For inputs whose
// buckets fit a block's LDS (uniform 2^28 keys: about 4096 rows each), one
// more read and write of the data sorts every bucket by the low digits in
// LDS, instead of one global pass per low digit: three trips through HBM
// instead of four for 32-bit keys.  A bucket larger than a block's capacity
// sets an overflow word; the host then finishes with ordinary passes (the
// rows with equal keys are still in input order, so stability holds).
constexpr uint32_t RB_BLOCK = 512;
constexpr uint32_t RB_ITEMS = 16;
constexpr uint32_t RB_CAP = RB_BLOCK * RB_ITEMS;   // rows one block sorts at once
constexpr uint32_t RB_NB = 1u << 16;               // buckets
// 32-bit keys (buckets = the high 16 bits): one LDS counting pass over
// RB_FBINS coarse bins (the bucket's offset in the chunk, then the top bits
// of the low 16), each bin then sorted by (low 16 bits, row in the chunk);
// a bin above RB_FIXMAX rows sends the chunk to the stable ranking instead
constexpr uint32_t RB_FBITS = 11;
constexpr uint32_t RB_FBINS = 1u << RB_FBITS;
constexpr uint32_t RB_FIXMAX = 32;

struct RbDesc {
    uint32_t nlow;                 // low digits sorted in LDS (1 or 2), lowest first
    uint32_t lshift[2], lmask[2];
    uint32_t sa, sb, mb;           // the two high digits (a is 8 bits wide; b's mask mb)
    uint32_t window;               // rows per block window
};

__device__ __forceinline__ uint32_t rb_bucket(uint32_t k, const RbDesc& d) {
    return ((k >> d.sa) & 0xFFu) | (((k >> d.sb) & d.mb) << 8);
}

// bstart[u] = first row of bucket u in rows sorted by (b, a); bstart[2^16]
// = n.  Block x = digit b (its run from the histogram of b), thread y =
// digit a: binary search of the run, whose rows are ordered by digit a.
__global__ __launch_bounds__(256) void k_rb_starts(const uint32_t* __restrict__ keys, uint32_t n,
                                                   const uint32_t* __restrict__ hist_b, RbDesc d,
                                                   uint32_t* __restrict__ bstart) {
    __shared__ uint32_t s_w[256 / PD_WAVE];
    const uint32_t t = threadIdx.x, db = blockIdx.x, l = lane(), w = t / PD_WAVE;
    const uint32_t h = t < db ? hist_b[t] : 0u;
    const uint32_t inc = wave_incl_scan(h);
    if (l == PD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t i = 0; i < 256 / PD_WAVE; i++) lo += s_w[i];
    uint32_t hi = lo + hist_b[db];
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (((keys[mid] >> d.sa) & 0xFFu) >= t) hi = mid;
        else lo = mid + 1;
    }
    bstart[db * 256 + t] = lo;
    if (db == 255 && t == 0) bstart[RB_NB] = n;
}

// Chunks of whole buckets, at most RB_CAP rows and 256 buckets each.
// Thread w takes the buckets that start in rows [w W, (w + 1) W) (W =
// window): one chunk when they fit, else chunks formed greedily bucket by
// bucket.  A bucket larger than RB_CAP sets *overflow and is left out.  A
// chunk is (first row, end row, first bucket << 8 | bits of its last
// bucket's offset from the first, 0 for one bucket);
// chunks are appended to `list` (one counter atomic per wave, after a
// counting walk and a wave scan).
__device__ __forceinline__ uint32_t rb_first_ge(const uint32_t* __restrict__ bstart, uint64_t x) {
    uint32_t lo = 0, hi = RB_NB;   // bstart is non-decreasing and bstart[2^16] = n >= x
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (bstart[mid] >= x) hi = mid;
        else lo = mid + 1;
    }
    return bstart[lo];
}

__global__ __launch_bounds__(256) void k_rb_chunks(const uint32_t* __restrict__ keys, uint32_t n, uint32_t nw,
                                                   const uint32_t* __restrict__ bstart, RbDesc d,
                                                   uint4* __restrict__ list, uint32_t* __restrict__ count,
                                                   uint32_t* __restrict__ overflow) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s = 0, e = 0;
    if (w < nw) {
        const uint64_t x0 = (uint64_t)w * d.window;
        s = rb_first_ge(bstart, x0);
        e = rb_first_ge(bstart, min((uint64_t)n, x0 + d.window));
    }
    // bits of a chunk's bucket offsets (0: one bucket)
    auto span_bits = [](uint32_t span) -> uint32_t { return span ? 32u - (uint32_t)__clz(span) : 0u; };
    // walk(emit): the chunks of [s, e); emit(first, end, info)
    auto walk = [&](auto&& emit) {
        if (s >= e) return;
        const uint32_t u0 = rb_bucket(keys[s], d), u1 = rb_bucket(keys[e - 1], d);
        if (e - s <= RB_CAP && u1 - u0 <= 255u) {
            emit(s, e, (u0 << 8) | span_bits(u1 - u0));
            return;
        }
        for (uint32_t p = s; p < e;) {
            const uint32_t b0 = rb_bucket(keys[p], d);
            const uint32_t e0 = bstart[b0 + 1];
            if (e0 - p > RB_CAP) {
                atomicOr(overflow, 1u);
                p = e0;
                continue;
            }
            const uint32_t lim = min(e, p + RB_CAP);
            uint32_t u = b0;
            while (u < b0 + 255u && u + 1 < RB_NB && bstart[u + 2] <= lim) u++;
            const uint32_t q = bstart[u + 1];
            emit(p, q, (b0 << 8) | (q > e0 ? span_bits(u - b0) : 0u));
            p = q;
        }
    };
    uint32_t cnt = 0;
    walk([&](uint32_t, uint32_t, uint32_t) { cnt++; });
    const uint32_t inc = wave_incl_scan(cnt);
    uint32_t base = 0;
    if (lane() == PD_WAVE - 1 && inc) base = atomicAdd(count, inc);
    base = __shfl(base, PD_WAVE - 1) + inc - cnt;
    walk([&](uint32_t a, uint32_t b, uint32_t info) { list[base++] = make_uint4(a, b, info, 0u); });
}

// Persistent: block b sorts chunks b, b + gridDim.x, ... of the list, in
// place: each chunk is loaded, ranked stably by each low digit in turn,
// then by bucket when it holds more than one, and written back (ranking by
// per-wave LDS match masks, as the onesweep pass ranks; one ballot per
// digit bit instead measured 2.22 against 1.68 ms at 2^28 pairs; 1024
// threads x 8 rows, 64 VGPRs with spills, 1.87 ms).  Padding lanes take the
// largest digit and rank after every real row.
template <typename V, bool HV>
__global__ __launch_bounds__(RB_BLOCK, 4) void k_rb_sort(uint32_t* __restrict__ keys, V* __restrict__ vals,
                                                        const uint4* __restrict__ list,
                                                        const uint32_t* __restrict__ count, RbDesc d) {
    constexpr uint32_t WAVES = RB_BLOCK / PD_WAVE;
    static_assert(RB_CAP * 4 >= WAVES * RS_BINS * 8, "match masks must fit in the key buffer");
    // 4-byte payloads travel with their keys as one 8-byte LDS word (one
    // LDS store per row at the scatter and one load at the reload and the
    // write-out, instead of two)
    constexpr bool PACK = HV && sizeof(V) == 4;
    __shared__ __align__(16) uint32_t s_keys[PACK ? 2 * RB_CAP : RB_CAP];
    __shared__ __align__(16) V s_vals[HV && !PACK ? RB_CAP : 1];
    uint2* s_kv = reinterpret_cast<uint2*>(s_keys);
    __shared__ uint16_t s_cnt[WAVES][RS_BINS];
    __shared__ uint32_t s_off[RS_BINS];
    __shared__ uint32_t s_wsum[RS_BINS / PD_WAVE];
    auto s_match = reinterpret_cast<uint64_t (*)[RS_BINS]>(s_keys);   // per-wave masks while ranking
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint64_t lt = lanemask_lt();
    const uint32_t nch = count[0];
    if (blockIdx.x >= nch) return;
    // rows per thread: a chunk is spread over all waves, wave w holding rows
    // [w 64 ipt, (w + 1) 64 ipt); items past the wave's rows are skipped
    // (wave-uniform), so only the last wave's last item holds padding lanes
    uint32_t key[RB_ITEMS];
    V val[HV ? RB_ITEMS : 1];
    uint32_t p = 0, m = 0, info = 0, wbase = 0, wrows = 0;
    auto load_chunk = [&](const uint4 c4) {
        p = (uint32_t)__builtin_amdgcn_readfirstlane(c4.x);
        m = (uint32_t)__builtin_amdgcn_readfirstlane(c4.y) - p;
        info = (uint32_t)__builtin_amdgcn_readfirstlane(c4.z);
        const uint32_t ipt = (m + RB_BLOCK - 1) / RB_BLOCK;
        wbase = w * (PD_WAVE * ipt);
        wrows = (uint32_t)__builtin_amdgcn_readfirstlane(m > wbase ? min(PD_WAVE * ipt, m - wbase) : 0u);
        const __amdgpu_buffer_rsrc_t rk = buf_rsrc(keys + p + wbase, wrows * 4u);
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++)
            key[i] = i * PD_WAVE < wrows ? __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0) : 0u;
        if constexpr (HV) {
            const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vals + p + wbase, wrows * (uint32_t)sizeof(V));
#pragma unroll
            for (uint32_t i = 0; i < RB_ITEMS; i++) {
                if (i * PD_WAVE >= wrows) val[i] = V(0);
                else if constexpr (sizeof(V) == 4)
                    val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, (l + i * PD_WAVE) * 4u, 0, 0);
                else
                    val[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rv, (l + i * PD_WAVE) * 8u, 0, 0));
            }
        }
    };
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        load_chunk(list[ci]);
        const uint32_t u0 = info >> 8;
        const uint32_t npass = d.nlow + ((info & 0xFFu) ? 1u : 0u);
        for (uint32_t j = 0; j < npass; j++) {
            const bool by_bucket = j >= d.nlow;
            const uint32_t sh = by_bucket ? 0u : d.lshift[j], mk = by_bucket ? 0u : d.lmask[j];
            const uint32_t dmax = RS_BINS - 1;   // padding lanes: the largest digit
            for (uint32_t i = l; i < RS_BINS; i += PD_WAVE) {
                s_match[w][i] = 0;
                s_cnt[w][i] = 0;
            }
            __syncthreads();
            uint32_t pos[RB_ITEMS];
#pragma unroll
            for (uint32_t i = 0; i < RB_ITEMS; i++) {
                if (i * PD_WAVE >= wrows) break;
                uint32_t kk = key[i];
                asm volatile("" : "+v"(kk));
                const uint32_t dg = (i * PD_WAVE + l) < wrows ? (by_bucket ? rb_bucket(kk, d) - u0 : (kk >> sh) & mk) : dmax;
                // lanes holding the same digit: their bits ORed into the
                // digit's LDS mask, read back
                atomicOr((unsigned long long*)&s_match[w][dg], (unsigned long long)(1ull << l));
                PD_COMPILER_BARRIER();
                const uint64_t mm = s_match[w][dg];
                const uint32_t old = s_cnt[w][dg];
                const uint32_t below = popc64(mm & lt);
                pos[i] = (old + below) | (dg << 16);
                PD_COMPILER_BARRIER();
                if (below == 0) {
                    s_cnt[w][dg] = (uint16_t)(old + popc64(mm));
                    s_match[w][dg] = 0;
                }
                PD_COMPILER_BARRIER();
            }
            __syncthreads();
            // per digit: prefix over waves, then over digits
            uint32_t cnt = 0, inc = 0;
            if (t < RS_BINS) {
#pragma unroll
                for (uint32_t ww = 0; ww < WAVES; ww++) {
                    const uint32_t c = s_cnt[ww][t];
                    s_cnt[ww][t] = (uint16_t)cnt;
                    cnt += c;
                }
                inc = wave_incl_scan(cnt);
                if (l == PD_WAVE - 1) s_wsum[w] = inc;
            }
            __syncthreads();
            if (t < RS_BINS) {
                uint32_t add = 0;
                for (uint32_t i = 0; i < w; i++) add += s_wsum[i];
                s_off[t] = add + inc - cnt;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t i = 0; i < RB_ITEMS; i++) {
                if (i * PD_WAVE >= wrows) break;
                const uint32_t dg = pos[i] >> 16;
                const uint32_t r = (pos[i] & 0xFFFFu) + s_off[dg] + s_cnt[w][dg];
                if constexpr (PACK) {
                    s_kv[r] = make_uint2(key[i], (uint32_t)val[i]);
                } else {
                    s_keys[r] = key[i];
                    if constexpr (HV) s_vals[r] = val[i];
                }
            }
            __syncthreads();
            if (j + 1 < npass) {
                // the next pass ranks the rows in their new order
#pragma unroll
                for (uint32_t i = 0; i < RB_ITEMS; i++) {
                    if (i * PD_WAVE >= wrows) break;
                    if constexpr (PACK) {
                        const uint2 e = s_kv[wbase + i * PD_WAVE + l];
                        key[i] = e.x;
                        val[i] = (V)e.y;
                    } else {
                        key[i] = s_keys[wbase + i * PD_WAVE + l];
                        if constexpr (HV) val[i] = s_vals[wbase + i * PD_WAVE + l];
                    }
                }
                __syncthreads();   // every row is read before the masks reuse s_keys
            }
        }
#pragma unroll 4
        for (uint32_t k = 0; k < RB_ITEMS; k++) {
            const uint32_t j = t + k * RB_BLOCK;
            if (j >= m) break;
            if constexpr (PACK) {
                const uint2 e = s_kv[j];
                keys[p + j] = e.x;
                vals[p + j] = (V)e.y;
            } else {
                keys[p + j] = s_keys[j];
                if constexpr (HV) vals[p + j] = s_vals[j];
            }
        }
        __syncthreads();   // the next chunk reuses the LDS buffers
    }
}

// Counting form of the bucket pass, for 32-bit keys whose buckets are the
// high 16 bits (four varying digits).  Persistent over the chunk list as
// k_rb_sort.  Every row takes a rank in its coarse bin (the bucket's offset
// in the chunk << tb | the low 16 bits >> (16 - tb), tb = RB_FBITS - bits
// of the chunk's bucket span: at most RB_FBINS bins) from one LDS atomic;
// a block scan gives the bins' starts; each row's sort word (low 16 bits
// << 16 | row in the chunk) is placed in its bin; a row's final place is
// its bin's start plus the words of the bin below its own (about three
// independent reads), which orders equal keys by input row, so the sort
// stays stable.  Rows then go from registers to their places in LDS and
// out in order.  One LDS atomic, one placement, a few reads and one store
// per row, against two ranking rounds of match masks.  A chunk with a bin
// above RB_FIXMAX rows is left untouched and listed in `retry` for
// k_rb_sort.
template <typename V, bool HV>
__global__ __launch_bounds__(RB_BLOCK, 4) void k_rb_count(uint32_t* __restrict__ keys, V* __restrict__ vals,
                                                         const uint4* __restrict__ list,
                                                         const uint32_t* __restrict__ count,
                                                         uint4* __restrict__ retry, uint32_t* __restrict__ nretry) {
    constexpr uint32_t WAVES = RB_BLOCK / PD_WAVE;
    constexpr bool PACK = HV && sizeof(V) == 4;
    static_assert(RB_FBINS == 4 * RB_BLOCK, "four bins per thread");
    static_assert(RB_CAP <= 65536, "rows in a chunk fit 16 bits");
    __shared__ __align__(16) uint32_t s_pk[PACK ? 2 * RB_CAP : RB_CAP];
    __shared__ __align__(16) V s_val[HV && !PACK ? RB_CAP : 1];
    __shared__ __align__(16) uint32_t s_bin[RB_FBINS + 4];   // counts, then starts; [RB_FBINS] = rows
    __shared__ uint32_t s_fw[WAVES];
    uint2* s_kv = reinterpret_cast<uint2*>(s_pk);
    auto e_put = [&](uint32_t r, uint32_t pk, V v) {
        if constexpr (PACK) {
            s_kv[r] = make_uint2(pk, (uint32_t)v);
        } else {
            s_pk[r] = pk;
            if constexpr (HV) s_val[r] = v;
        }
    };
    auto e_pk = [&](uint32_t r) -> uint32_t {
        if constexpr (PACK) return s_kv[r].x;
        else return s_pk[r];
    };
    auto e_val = [&](uint32_t r) -> V {
        if constexpr (PACK) return (V)s_kv[r].y;
        else if constexpr (HV) return s_val[r];
        else return V(0);
    };
    const uint32_t t = threadIdx.x, l = lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    const uint32_t nch = count[0];
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        const uint4 c4 = list[ci];
        const uint32_t p = (uint32_t)__builtin_amdgcn_readfirstlane(c4.x);
        const uint32_t m = (uint32_t)__builtin_amdgcn_readfirstlane(c4.y) - p;
        const uint32_t info = (uint32_t)__builtin_amdgcn_readfirstlane(c4.z);
        const uint32_t u0 = info >> 8, tb = RB_FBITS - (info & 0xFFu);
        // wave w holds rows [w 64 ipt, (w + 1) 64 ipt), as k_rb_sort
        const uint32_t ipt = (m + RB_BLOCK - 1) / RB_BLOCK;
        const uint32_t wbase = w * (PD_WAVE * ipt);
        const uint32_t wrows = (uint32_t)__builtin_amdgcn_readfirstlane(m > wbase ? min(PD_WAVE * ipt, m - wbase) : 0u);
        for (uint32_t i = t; i < RB_FBINS + 4; i += RB_BLOCK) s_bin[i] = 0;
        const __amdgpu_buffer_rsrc_t rk = buf_rsrc(keys + p + wbase, wrows * 4u);
        uint32_t key[RB_ITEMS];
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++)
            key[i] = i * PD_WAVE < wrows ? __builtin_amdgcn_raw_buffer_load_b32(rk, (l + i * PD_WAVE) * 4u, 0, 0) : 0u;
        V val[HV ? RB_ITEMS : 1];
        if constexpr (HV) {
            const __amdgpu_buffer_rsrc_t rv = buf_rsrc(vals + p + wbase, wrows * (uint32_t)sizeof(V));
#pragma unroll
            for (uint32_t i = 0; i < RB_ITEMS; i++) {
                if (i * PD_WAVE >= wrows) val[i] = V(0);
                else if constexpr (sizeof(V) == 4)
                    val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, (l + i * PD_WAVE) * 4u, 0, 0);
                else
                    val[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rv, (l + i * PD_WAVE) * 8u, 0, 0));
            }
        }
        __syncthreads();
        uint32_t br[RB_ITEMS];   // bin << 16 | rank in the bin
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++) {
            if (i * PD_WAVE >= wrows) break;
            const uint32_t b = (((key[i] >> 16) - u0) << tb) | ((key[i] & 0xFFFFu) >> (16 - tb));
            if (i * PD_WAVE + l < wrows) br[i] = (b << 16) | atomicAdd(&s_bin[b], 1u);
        }
        __syncthreads();
        // exclusive scan of the counts: thread t holds bins [4t, 4t + 4)
        const uint4 n4 = reinterpret_cast<const uint4*>(s_bin)[t];
        const uint32_t tsum = n4.x + n4.y + n4.z + n4.w;
        const uint32_t tinc = wave_incl_scan(tsum);
        if (l == PD_WAVE - 1) s_fw[w] = tinc;
        __syncthreads();
        uint32_t st = tinc - tsum;
        for (uint32_t i = 0; i < w; i++) st += s_fw[i];
        reinterpret_cast<uint4*>(s_bin)[t] = make_uint4(st, st + n4.x, st + n4.x + n4.y, st + n4.x + n4.y + n4.z);
        if (t == RB_BLOCK - 1) s_bin[RB_FBINS] = st + tsum;
        __syncthreads();
        // the sort words (low 16 bits << 16 | row) placed by bin
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++) {
            if (i * PD_WAVE >= wrows) break;
            if (i * PD_WAVE + l < wrows) {
                const uint32_t b = br[i] >> 16;
                const uint32_t b0 = s_bin[b], b1 = s_bin[b + 1];
                s_pk[b0 + (br[i] & 0xFFFFu)] = (key[i] << 16) | (wbase + i * PD_WAVE + l);
                br[i] = (b0 << 16) | (b1 - b0);   // the bin's start and size
            }
        }
        __syncthreads();
        // a row's place: its bin's start + the bin's words below its own
        // (one pass over a bin of about three words per row; reading word q
        // of every row's bin in one step took 1.66 ms against 1.38 ms, since
        // every step then issues a read for each item)
        bool big = false;
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++) {
            if (i * PD_WAVE >= wrows) break;
            if (i * PD_WAVE + l < wrows) {
                const uint32_t b0 = br[i] >> 16, c = br[i] & 0xFFFFu;
                const uint32_t mine = (key[i] << 16) | (wbase + i * PD_WAVE + l);
                uint32_t below = 0;
                if (c > RB_FIXMAX) {
                    big = true;
                } else {
                    for (uint32_t q = 0; q < c; q++) below += s_pk[b0 + q] < mine ? 1u : 0u;
                }
                br[i] = b0 + below;
            }
        }
        if (__syncthreads_or(big)) {
            if (t == 0) retry[atomicAdd(nretry, 1u)] = c4;
            continue;   // the chunk is untouched; the barrier above ended every LDS read
        }
#pragma unroll
        for (uint32_t i = 0; i < RB_ITEMS; i++) {
            if (i * PD_WAVE >= wrows) break;
            if (i * PD_WAVE + l < wrows) e_put(br[i], key[i], val[HV ? i : 0]);
        }
        __syncthreads();
#pragma unroll 4
        for (uint32_t k = 0; k < RB_ITEMS; k++) {
            const uint32_t j = t + k * RB_BLOCK;
            if (j >= m) break;
            keys[p + j] = e_pk(j);
            if constexpr (HV) vals[p + j] = e_val(j);
        }
        __syncthreads();   // the next chunk reuses the LDS buffers
    }
}

// Bucket scheme for the digits of `run` (knob sort_bucket).  The bucket
// pass is worth one global pass only when it replaces two (four digits) and
// its chunks are large: at 2^28 uniform pairs (buckets of about 4096 rows)
// it takes 1.67 ms where two onesweep passes take 2.06 ms; at 2^27 (2048-row
// buckets, two per chunk, a third LDS ranking) 1.31 ms against 1.06 ms.  So
// by default (1) it runs for four varying digits and a largest bucket,
// estimated from the histograms of the two high digits as if independent,
// of 3072 to 3/4 of a block's capacity rows; 2 takes it for any 3 or 4
// digits whose estimate fits (tests); 0 never.  `hh` holds the two
// histograms.
static bool use_bucket(pimdal_ctx* c, const RsPassDesc& run, uint64_t n, const uint32_t* hh) {
    const int64_t mode = c->tune_get("sort_bucket", 1);
    if (mode == 0 || c->capture || run.npass < 3) return false;
    if (run.mask[run.npass - 2] != 0xFFu) return false;
    uint32_t ma = 0, mb = 0;
    for (int i = 0; i < RS_BINS; i++) {
        ma = std::max(ma, hh[i]);
        mb = std::max(mb, hh[RS_BINS + i]);
    }
    const double est = (double)ma * (double)mb / (double)n;
    if (est > 0.75 * RB_CAP) return false;
    return mode == 2 || (run.npass == 4 && est >= 3072.0);
}

template <typename V, bool HV>
static void sort_bucket(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                        const RsPassDesc& run, uint32_t* hist, uint32_t* offs, uint32_t* ctrs, bool low_counted,
                        hipStream_t s) {
    const uint32_t P = run.npass;
    // the two high digits: in -> scratch -> out
    RsPassDesc hi{};
    hi.npass = 2;
    for (int j = 0; j < 2; j++) {
        hi.shift[j] = run.shift[P - 2 + j];
        hi.mask[j] = run.mask[P - 2 + j];
    }
    run_passes<V, HV>(c, kin, vin, n, kout, vout, hi, hist + (P - 2) * RS_BINS, offs + (P - 2) * RS_BINS,
                      ctrs + (P - 2), s);
    RbDesc d{};
    d.nlow = P - 2;
    for (uint32_t j = 0; j < d.nlow; j++) {
        d.lshift[j] = run.shift[j];
        d.lmask[j] = run.mask[j];
    }
    d.sa = run.shift[P - 2];
    d.sb = run.shift[P - 1];
    d.mb = run.mask[P - 1];
    d.window = RB_CAP / 2;
    const uint32_t nw = (uint32_t)ceil_div(n, (uint64_t)d.window);
    // bstart[2^16 + 1] | overflow | chunk count | chunk list (at most one
    // chunk per window plus one per bucket)
    // the counting form needs 32-bit keys bucketed by their high 16 bits
    // (knob sort_bucket_count, default 1)
    const bool counting = c->tune_get("sort_bucket_count", 1) != 0 && d.nlow == 2 && d.lshift[0] == 0 &&
                          d.lshift[1] == 8 && d.lmask[0] == 0xFFu && d.lmask[1] == 0xFFu && d.sa == 16 &&
                          d.sb == 24 && d.mb == 0xFFu;
    // bstart[2^16 + 1] | overflow | chunk count | retry count | chunk list
    // (at most one chunk per window plus one per bucket) | retry list
    const uint64_t maxch = (uint64_t)nw + RB_NB;
    const uint64_t head = (RB_NB + 1 + 3) * 4;
    Scratch sc(c, head + maxch * 16 * (counting ? 2 : 1), s);
    uint32_t* bstart = sc.as<uint32_t>();
    uint32_t* ovf = bstart + RB_NB + 1;
    uint32_t* cnt = ovf + 1;
    uint32_t* nretry = ovf + 2;
    uint4* list = reinterpret_cast<uint4*>(sc.as<char>() + head);
    uint4* retry = list + maxch;
    fill_u32(c, ovf, 0, 3, s);
    launch(c, "radix_bstart", k_rb_starts, dim3(256), dim3(256), 0, s, (const uint32_t*)kout, (uint32_t)n,
           hist + (P - 1) * RS_BINS, d, bstart);
    launch(c, "radix_chunks", k_rb_chunks, dim3((nw + 255) / 256), dim3(256), 0, s, (const uint32_t*)kout, (uint32_t)n,
           nw, (const uint32_t*)bstart, d, list, cnt, ovf);
    const uint32_t grid = std::max<uint32_t>(1, blocks_per_cu(k_rb_sort<V, HV>, RB_BLOCK) * c->cu_count);
    if (counting) {
        // chunks with a large bin (many equal low bits) go to the ranking
        const uint32_t cgrid = std::max<uint32_t>(1, blocks_per_cu(k_rb_count<V, HV>, RB_BLOCK) * c->cu_count);
        launch(c, "radix_bucket", k_rb_count<V, HV>, dim3(cgrid), dim3(RB_BLOCK), 0, s, kout, vout,
               (const uint4*)list, (const uint32_t*)cnt, retry, nretry);
        launch(c, "radix_bucket_retry", k_rb_sort<V, HV>, dim3(grid), dim3(RB_BLOCK), 0, s, kout, vout,
               (const uint4*)retry, (const uint32_t*)nretry, d);
    } else {
        launch(c, "radix_bucket", k_rb_sort<V, HV>, dim3(grid), dim3(RB_BLOCK), 0, s, kout, vout, (const uint4*)list,
               (const uint32_t*)cnt, d);
    }
    uint32_t of = 0;
    read_small(c, &of, ovf, 4, s);
    if (of) {
        // a bucket did not fit: every digit again, by ordinary passes
        Scratch tk(c, n * 4, s), tv(c, HV ? n * sizeof(V) : 0, s);
        copy_d2d(c, tk.get(), kout, n * 4, s);
        if (HV) copy_d2d(c, tv.get(), vout, n * sizeof(V), s);
        fill_u32(c, ctrs, 0, 4, s);
        if (!low_counted) {
            // only the high digits were counted: the low ones now (four
            // digits at shifts 0, 8, 16, 24)
            RsPassDesc lo2{};
            lo2.npass = 2;
            for (int j = 0; j < 2; j++) {
                lo2.shift[j] = 8 * j;
                lo2.mask[j] = 0xFF;
            }
            const uint32_t hg = (uint32_t)std::min<uint64_t>(ceil_div(n, 512 * 16), (uint64_t)c->cu_count * 4);
            // (ctrs - 2: the AND/OR words, which the same keys leave as they are)
            launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hg)), dim3(512), 0, s,
                   (const uint32_t*)tk.as<uint32_t>(), n, lo2, hist, ctrs - 2);
        }
        run_passes<V, HV>(c, tk.as<uint32_t>(), HV ? tv.as<V>() : (const V*)nullptr, n, kout, vout, run, hist, offs,
                          ctrs, s);
    }
}

void sort_pairs(pimdal_ctx* c, const uint32_t* kin, const void* vin, uint32_t vw, uint64_t n, uint32_t* kout,
                void* vout, int begin_bit, int end_bit, hipStream_t s) {
    PD_REQUIRE(vw == 0 || vw == 4 || vw == 8, PIMDAL_E_UNSUPPORTED, "payload width must be 0, 4 or 8 bytes");
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "radix sort supports up to 2^30-1 pairs per call");
    PD_REQUIRE(n == 0 || (kin && kout), PIMDAL_E_INVALID, "key pointer is NULL");
    PD_REQUIRE(vw == 0 || n == 0 || (vin && vout), PIMDAL_E_INVALID, "payload pointer is NULL");
    PD_REQUIRE(n == 0 || kin != kout, PIMDAL_E_INVALID, "in-place sort is not supported (keys_in == keys_out)");
    PD_REQUIRE(vw == 0 || n == 0 || vin != vout, PIMDAL_E_INVALID, "in-place sort is not supported (vals_in == vals_out)");
    const bool detect = begin_bit < 0;
    PD_REQUIRE(detect || (begin_bit >= 0 && begin_bit <= end_bit && end_bit <= 32), PIMDAL_E_INVALID,
               "bad bit range");
    if (n == 0) return;

    RsPassDesc all{};
    if (detect) {
        all.npass = 4;
        for (int p = 0; p < 4; p++) { all.shift[p] = 8 * p; all.mask[p] = 0xFF; }
    } else {
        all.npass = 0;
        for (int b = begin_bit; b < end_bit; b += 8) {
            all.shift[all.npass] = b;
            all.mask[all.npass] = (uint32_t)((1ull << std::min(8, end_bit - b)) - 1);
            all.npass++;
        }
    }
    if (use_rts(c)) {
        if (vw == 0)
            sort_rts<uint32_t, false>(c, kin, nullptr, n, kout, nullptr, detect, all, s);
        else if (vw == 4)
            sort_rts<uint32_t, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, detect, all, s);
        else
            sort_rts<uint64_t, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, detect, all, s);
        return;
    }

    // chained passes: one histogram read of the keys counts every digit
    // small scratch: hist[4*256] | offs[4*256] | andor[2] | ctr[4]
    Scratch small(c, (4 * RS_BINS * 2 + 8) * 4, s);
    uint32_t* hist = small.as<uint32_t>();
    uint32_t* offs = hist + 4 * RS_BINS;
    uint32_t* andor = offs + 4 * RS_BINS;
    uint32_t* ctrs = andor + 2;
    fill_u32(c, hist, 0, 4 * RS_BINS * 2, s);
    fill_u32(c, andor, 0xFFFFFFFFu, 1, s);
    fill_u32(c, andor + 1, 0, 1 + 4, s);
    const uint32_t hist_grid = (uint32_t)std::min<uint64_t>(ceil_div(n, 512 * 16), (uint64_t)c->cu_count * 4);
    const bool bucket_ok = c->tune_get("sort_bucket", 1) != 0 && !c->capture;
    // Where the bucket scheme is all but certain (four digits, buckets of
    // at least n / 2^16 >= 3072 rows), the first read counts only the two
    // high digits, which are all it uses; the low digits are counted later
    // if ordinary passes turn out to need them.
    bool std4 = all.npass == 4;
    for (int p = 0; p < 4 && std4; p++) std4 = all.shift[p] == 8u * p && all.mask[p] == 0xFFu;
    const bool hi_first = bucket_ok && std4 && c->tune_get("sort_bucket", 1) == 1 && n >= 3072ull * RB_NB;
    RsPassDesc lo2{}, hi2{};
    lo2.npass = hi2.npass = 2;
    for (int j = 0; j < 2; j++) {
        lo2.shift[j] = 8 * j;
        hi2.shift[j] = 16 + 8 * j;
        lo2.mask[j] = hi2.mask[j] = 0xFF;
    }
    auto count_low = [&](const uint32_t* keys) {
        launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, keys, n, lo2, hist,
               andor);
    };
    if (hi_first)
        launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, kin, n, hi2,
               hist + 2 * RS_BINS, andor);
    else if (all.npass)
        launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, kin, n, all, hist, andor);

    RsPassDesc run = all;
    // host copy of hist[4][256] | offs | andor, read once when the digit
    // detection or the bucket scheme's check needs it
    std::vector<uint32_t> hh;
    if (detect || (bucket_ok && all.npass >= 3)) {
        hh.resize(4 * RS_BINS * 2 + 2);
        read_small(c, hh.data(), hist, hh.size() * 4, s);
    }
    int orig[4] = {0, 1, 2, 3};   // run digit -> its histogram in hh
    if (hi_first) {
        // the bucket scheme now, or the low digits counted for ordinary passes
        // with the digits detected, all four must vary
        const uint32_t varying = hh[4 * RS_BINS * 2] ^ hh[4 * RS_BINS * 2 + 1];
        bool four = true;
        for (int p = 0; p < 4; p++) four = four && (!detect || ((varying >> (8 * p)) & 0xFF) != 0);
        if (four && use_bucket(c, all, n, hh.data() + 2 * RS_BINS)) {
            if (vw == 0)
                sort_bucket<uint32_t, false>(c, kin, nullptr, n, kout, nullptr, all, hist, offs, ctrs, false, s);
            else if (vw == 4)
                sort_bucket<uint32_t, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, all, hist, offs,
                                            ctrs, false, s);
            else
                sort_bucket<uint64_t, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, all, hist, offs,
                                            ctrs, false, s);
            return;
        }
        count_low(kin);
        read_small(c, hh.data(), hist, 2 * RS_BINS * 4, s);
    }
    if (detect) {
        const uint32_t varying = hh[4 * RS_BINS * 2] ^ hh[4 * RS_BINS * 2 + 1];
        // keep only digits with a varying bit; move their histograms down
        run.npass = 0;
        for (int p = 0; p < 4; p++) {
            if ((varying >> (8 * p)) & 0xFF) {
                if (run.npass != (uint32_t)p)
                    copy_d2d(c, hist + run.npass * RS_BINS, hist + p * RS_BINS, RS_BINS * 4, s);
                orig[run.npass] = p;
                run.shift[run.npass] = 8 * p;
                run.mask[run.npass] = 0xFF;
                run.npass++;
            }
        }
    }
    if (run.npass == 0) {
        copy_d2d(c, kout, kin, n * 4, s);
        if (vw) copy_d2d(c, vout, vin, n * vw, s);
        return;
    }
    if (bucket_ok && run.npass >= 3) {
        std::vector<uint32_t> h2(2 * RS_BINS);
        std::copy_n(hh.begin() + orig[run.npass - 2] * RS_BINS, RS_BINS, h2.begin());
        std::copy_n(hh.begin() + orig[run.npass - 1] * RS_BINS, RS_BINS, h2.begin() + RS_BINS);
        if (use_bucket(c, run, n, h2.data())) {
            if (vw == 0)
                sort_bucket<uint32_t, false>(c, kin, nullptr, n, kout, nullptr, run, hist, offs, ctrs, true, s);
            else if (vw == 4)
                sort_bucket<uint32_t, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, hist, offs,
                                            ctrs, true, s);
            else
                sort_bucket<uint64_t, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, hist, offs,
                                            ctrs, true, s);
            return;
        }
    }
    if (vw == 0)
        run_passes<uint32_t, false>(c, kin, nullptr, n, kout, nullptr, run, hist, offs, ctrs, s);
    else if (vw == 4)
        run_passes<uint32_t, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, hist, offs, ctrs, s);
    else
        run_passes<uint64_t, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, hist, offs, ctrs, s);
}

// Histogram of the whole partition id ref_hash(key) & (F - 1) for F <= 2^15
// (one LDS atomic per key over F bins, so lanes of an instruction rarely
// collide), folded in the same block into the per-pass digit histograms,
// which are added to the global ones (npass x 256 atomics per block).
// MAXB = 14: 64 KB of counters, 512 threads; MAXB = 15: 128 KB, 1024
// threads (one block per CU, 16 waves).
constexpr uint32_t RS_FULL_MAX_BITS = 15;
template <uint32_t MAXB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_part_hist_full(const uint32_t* __restrict__ keys, uint64_t n, uint32_t F,
                                                          RsPassDesc pd, uint32_t* __restrict__ g_hist) {
    __shared__ uint32_t h[1u << MAXB];
    __shared__ uint32_t hd[4][RS_BINS];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < F; i += blockDim.x) h[i] = 0;
    for (uint32_t i = t; i < 4 * RS_BINS; i += blockDim.x) (&hd[0][0])[i] = 0;
    __syncthreads();
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t r0 = min(n, (uint64_t)blockIdx.x * chunk), r1 = min(n, r0 + chunk);
    // 16-byte loads over the 4-aligned middle of the block's rows
    const uint64_t a0 = min(r1, (r0 + 3) & ~3ull), a1 = max(a0, r1 & ~3ull);
    const bool vec = ((uintptr_t)keys % 16) == 0;
    const uint64_t v1 = vec ? a1 / 4 : a0 / 4;
    for (uint64_t i = r0 + t; i < (vec ? a0 : r0); i += blockDim.x) atomicAdd(&h[pd::ref_hash(keys[i]) & (F - 1)], 1u);
    auto count4 = [&](const uint4 q) {
        atomicAdd(&h[pd::ref_hash(q.x) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.y) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.z) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.w) & (F - 1)], 1u);
    };
    // four 16-byte loads in flight per thread, then their atomics
    const uint4* k4p = reinterpret_cast<const uint4*>(keys);
    uint64_t v = a0 / 4 + t;
    for (; v + 3 * (uint64_t)BLOCK < v1; v += 4 * (uint64_t)BLOCK) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = k4p[v + (uint64_t)u * BLOCK];
#pragma unroll
        for (int u = 0; u < 4; u++) count4(q[u]);
    }
    for (; v < v1; v += BLOCK) count4(k4p[v]);
    for (uint64_t i = (vec ? a1 : r0) + t; i < r1; i += blockDim.x) atomicAdd(&h[pd::ref_hash(keys[i]) & (F - 1)], 1u);
    __syncthreads();
    for (uint32_t b = t; b < F; b += blockDim.x) {
        const uint32_t v = h[b];
        if (v)
            for (uint32_t p = 0; p < pd.npass; p++) atomicAdd(&hd[p][(b >> pd.shift[p]) & pd.mask[p]], v);
    }
    __syncthreads();
    for (uint32_t i = t; i < pd.npass * RS_BINS; i += blockDim.x) {
        const uint32_t v = (&hd[0][0])[i];
        if (v) atomicAdd(&g_hist[i], v);
    }
}

// Partition bounds of rows stably sorted by ref_hash(key) & (F - 1):
// bounds[f] = first row whose partition is >= f (binary search), bounds[F] = n.
__global__ void k_part_bounds(const uint32_t* __restrict__ k, uint64_t n, uint32_t F, uint32_t* __restrict__ bounds) {
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f <= F; f += gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;   // answer in [lo, hi]
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if ((pd::ref_hash(k[mid]) & (F - 1)) < f) lo = mid + 1; else hi = mid;
        }
        bounds[f] = (uint32_t)(f == F ? n : lo);
    }
}

void hash_partition(pimdal_ctx* c, const uint32_t* kin, const void* vin, uint32_t vw, uint64_t n, uint32_t bits,
                    uint32_t* kout, void* vout, uint32_t* bounds, hipStream_t s) {
    PD_REQUIRE(vw == 0 || vw == 4 || vw == 8, PIMDAL_E_UNSUPPORTED, "partition payload must be 0, 4 or 8 bytes");
    PD_REQUIRE(bits >= 1 && bits <= 24, PIMDAL_E_INVALID, "partition bits must be in [1, 24]");
    PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "hash partition supports up to 2^30-1 rows per call");
    const uint32_t F = 1u << bits;
    if (n == 0) {
        fill_u32(c, bounds, 0, F + 1, s);
        return;
    }
    // passes low digit first, widths as equal as possible (longer write runs)
    RsPassDesc run{};
    run.npass = (bits + 7) / 8;
    for (uint32_t p = 0, b = 0; p < run.npass; p++) {
        const uint32_t w = (bits - b) / (run.npass - p);
        run.shift[p] = b;
        run.mask[p] = (1u << w) - 1;
        b += w;
    }
    if (use_rts(c)) {
        RtsBufs B(c, n, false, s);
        if (vw == 4)
            run_passes_rts<uint32_t, true, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, B, false, s);
        else if (vw == 8)
            run_passes_rts<uint64_t, true, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, B, false, s);
        else
            run_passes_rts<uint32_t, false, true>(c, kin, nullptr, n, kout, nullptr, run, B, false, s);
        launch(c, "part_bounds", k_part_bounds, dim3((uint32_t)std::min<uint64_t>(ceil_div(F + 1, 256), 1024)), dim3(256),
               0, s, (const uint32_t*)kout, n, F, bounds);
        return;
    }
    // chained passes: every pass's digit histogram from one read of the keys
    Scratch small(c, (4 * RS_BINS * 2 + 8) * 4, s);
    uint32_t* hist = small.as<uint32_t>();
    uint32_t* offs = hist + 4 * RS_BINS;
    uint32_t* andor = offs + 4 * RS_BINS;
    uint32_t* ctrs = andor + 2;
    fill_u32(c, hist, 0, 4 * RS_BINS * 2 + 8, s);
    const uint32_t hist_grid = (uint32_t)std::min<uint64_t>(ceil_div(n, 512 * 16), (uint64_t)c->cu_count * 4);
    if (bits <= RS_FULL_MAX_BITS) {
        if (bits <= 14) {
            const uint32_t G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 16384), (uint64_t)c->cu_count * 2));
            launch(c, "part_hist", k_part_hist_full<14, 512>, dim3(G), dim3(512), 0, s, kin, n, F, run, hist);
        } else {
            const uint32_t G = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 32768), (uint64_t)c->cu_count));
            launch(c, "part_hist", k_part_hist_full<15, 1024>, dim3(G), dim3(1024), 0, s, kin, n, F, run, hist);
        }
    } else {
        launch(c, "part_hist", k_rs_hist<true>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, kin, n, run, hist, andor);
    }
    if (vw == 4)
        run_passes<uint32_t, true, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, hist, offs, ctrs, s);
    else if (vw == 8)
        run_passes<uint64_t, true, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, hist, offs, ctrs, s);
    else
        run_passes<uint32_t, false, true>(c, kin, nullptr, n, kout, nullptr, run, hist, offs, ctrs, s);
    launch(c, "part_bounds", k_part_bounds, dim3((uint32_t)std::min<uint64_t>(ceil_div(F + 1, 256), 1024)), dim3(256), 0,
           s, (const uint32_t*)kout, n, F, bounds);
}

// The last pass of hash_partition over rows already partitioned by hash
// bits [0, lo_bits) (in any order inside those partitions): one stable
// onesweep pass over bits [lo_bits, bits) from that digit's histogram
// (hist_hi, RS_BINS words), then the partition bounds.
void hash_partition_finish(pimdal_ctx* c, const uint32_t* kin, const uint64_t* vin, uint64_t n, uint32_t bits,
                           uint32_t lo_bits, const uint32_t* hist_hi, uint32_t* kout, uint64_t* vout, uint32_t* bounds,
                           hipStream_t s) {
    PD_REQUIRE(lo_bits >= 1 && bits > lo_bits && bits - lo_bits <= 8 && bits <= 24, PIMDAL_E_INVALID,
               "partition finish: one digit of at most 8 bits above the low bits");
    PD_REQUIRE(n >= 1 && n < (1ull << 30), PIMDAL_E_INVALID, "partition finish: 1 to 2^30-1 rows");
    const uint32_t F = 1u << bits;
    RsPassDesc run{};
    run.npass = 1;
    run.shift[0] = lo_bits;
    run.mask[0] = (1u << (bits - lo_bits)) - 1;
    Scratch small(c, (RS_BINS + 8) * 4, s);   // offs | ctrs
    uint32_t* offs = small.as<uint32_t>();
    uint32_t* ctrs = offs + RS_BINS;
    fill_u32(c, ctrs, 0, 8, s);
    run_passes<uint64_t, true, true>(c, kin, vin, n, kout, vout, run, hist_hi, offs, ctrs, s);
    launch(c, "part_bounds", k_part_bounds, dim3((uint32_t)std::min<uint64_t>(ceil_div(F + 1, 256), 1024)), dim3(256), 0,
           s, (const uint32_t*)kout, n, F, bounds);
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_sort_pairs_dev(pimdal_ctx* c, const uint32_t* keys_in, const void* vals_in, uint32_t val_width,
                                     uint64_t n, uint32_t* keys_out, void* vals_out, int begin_bit, int end_bit) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "ctx is NULL");
    sort_pairs(c, keys_in, vals_in, val_width, n, keys_out, vals_out, begin_bit, end_bit, c->stream);
    PD_API_END
}

extern "C" int pimdal_sort_pairs(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* payload,
                                 pimdal_col** out_key, pimdal_col** out_payload) {
    PD_API_BEGIN
    PD_REQUIRE(c && key && out_key, PIMDAL_E_INVALID, "NULL argument");
    PD_NO_CAPTURE(c);
    PD_REQUIRE(key->type == PIMDAL_U32, PIMDAL_E_UNSUPPORTED, "sort key must be uint32");
    PD_REQUIRE(!payload || payload->length == key->length, PIMDAL_E_INVALID, "payload length differs");
    PD_REQUIRE(!payload || out_payload, PIMDAL_E_INVALID, "out_payload is NULL");
    const uint64_t n = key->length;
    const uint32_t vw = payload ? type_width(payload->type) : 0;
    PD_REQUIRE(vw == 0 || vw == 4 || vw == 8, PIMDAL_E_UNSUPPORTED, "payload must be 4 or 8 bytes wide");
    pimdal_col* ok = nullptr;
    pimdal_col* ov = nullptr;
    try {
        ok = new pimdal_col{PIMDAL_U32, n, n, c->pool.alloc(std::max<uint64_t>(n, 1) * 4, c->stream), true};
        if (payload) ov = new pimdal_col{payload->type, n, n, c->pool.alloc(std::max<uint64_t>(n, 1) * vw, c->stream), true};
        sort_pairs(c, (const uint32_t*)key->data, payload ? payload->data : nullptr, vw, n, (uint32_t*)ok->data,
                   ov ? ov->data : nullptr, -1, -1, c->stream);
        PD_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        if (ok) { c->pool.free(ok->data, c->stream); delete ok; }
        if (ov) { c->pool.free(ov->data, c->stream); delete ov; }
        throw;
    }
    *out_key = ok;
    if (out_payload) *out_payload = ov;
    PD_API_END
}
