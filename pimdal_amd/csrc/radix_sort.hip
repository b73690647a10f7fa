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
    // four 16-byte loads in flight per thread, then their atomics
    uint64_t v = blockIdx.x * (uint64_t)blockDim.x + t;
    for (; v + 3 * stride < nvec; v += 4 * stride) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = k4p[v + u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++) count4(q[u]);
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
    // wave AND/OR, one atomic per wave
#pragma unroll
    for (int d = PD_WAVE / 2; d > 0; d >>= 1) {
        a &= __shfl_xor(a, d, PD_WAVE);
        o |= __shfl_xor(o, d, PD_WAVE);
    }
    if (pd::lane() == 0) {
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

// ---- one onesweep pass ----
// Compiler-only barrier: LDS instructions of one wave execute in issue order,
// so the ranking steps below need ordering against code motion, not waits.
// (Plain LDS accesses between these barriers, not volatile ones: the
// compiler follows a volatile access with a wait for every outstanding
// global load and store too, which would stall the ranking on the payload
// loads and, in the persistent pass, on the previous tile's stores.)
#define PD_COMPILER_BARRIER() __asm__ volatile("" ::: "memory")

// 4 waves per SIMD: at most 128 VGPRs (second launch-bounds argument is the
// minimum number of waves per execution unit).  EARLY: resolve the global
// offsets (look-back) before the LDS reorder instead of after it.
// BALLOT: rank with per-bit ballots instead of LDS match masks.
// LBW: statuses each digit's thread reads per look-back round trip.  A
// status written on another XCD is read from memory (the L2s are per XCD),
// so one round trip costs about a microsecond; with tiles finishing at tens
// per microsecond, a tile typically walks back over dozens of published
// aggregates before it reaches an inclusive prefix, and a wide window
// covers them in one or two round trips.
template <typename V, bool HAS_VALS, int BLOCK, bool EARLY, bool HASH, bool BALLOT = false, int LBW = 4,
          int ITEMS = RS_ITEMS>
__global__ __launch_bounds__(BLOCK, ITEMS <= 8 ? 8 : 4) void k_rs_onesweep(
    const uint32_t* __restrict__ kin, const V* __restrict__ vin, uint32_t* __restrict__ kout, V* __restrict__ vout,
    uint64_t n, uint32_t shift, uint32_t mask, const uint32_t* __restrict__ offs, uint32_t* __restrict__ status,
    uint32_t* __restrict__ tile_ctr) {
    constexpr int WAVES = BLOCK / PD_WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(TILE * 4 >= WAVES * RS_BINS * 8, "match masks must fit in the key buffer");
    __shared__ uint32_t s_tile;
    // keys reordered by digit; the per-wave match masks live here during ranking
    __shared__ __align__(16) uint32_t s_keys[TILE];
    __shared__ __align__(16) V s_vals[HAS_VALS ? TILE : 1];
    __shared__ uint16_t s_cnt[WAVES][RS_BINS];
    __shared__ uint32_t s_off[RS_BINS];   // tile-local digit start
    __shared__ uint32_t s_gbase[RS_BINS]; // global position of the digit's run minus its local start
    __shared__ uint32_t s_wsum[RS_BINS / PD_WAVE];
    auto s_match = reinterpret_cast<uint64_t (*)[RS_BINS]>(s_keys);

    const uint32_t t = threadIdx.x, l = pd::lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t / PD_WAVE);
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (uint32_t i = l; i < RS_BINS; i += PD_WAVE) {
        s_match[w][i] = 0;
        s_cnt[w][i] = 0;
    }
    __syncthreads();
    const uint64_t tile = s_tile;
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
        key[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, (l + (uint32_t)i * PD_WAVE) * 4u, 0, 0);
    if constexpr (HAS_VALS) {
        const __amdgpu_buffer_rsrc_t rv = pd::buf_rsrc(vin + wave_base, wrows * (uint32_t)sizeof(V));
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            if constexpr (sizeof(V) == 4)
                val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, (l + (uint32_t)i * PD_WAVE) * 4u, 0, 0);
            else
                val[i] = __builtin_bit_cast(
                    V, __builtin_amdgcn_raw_buffer_load_b64(rv, (l + (uint32_t)i * PD_WAVE) * 8u, 0, 0));
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

    // ---- per digit: prefix over waves, tile count, publish the aggregate ----
    uint32_t cnt_d = 0, inc_d = 0, excl = 0;
    if (t < RS_BINS) {
#pragma unroll
        for (int ww = 0; ww < WAVES; ww++) {
            const uint32_t c = s_cnt[ww][t];
            s_cnt[ww][t] = (uint16_t)cnt_d;
            cnt_d += c;
        }
        pd::st_relaxed(&status[tile * RS_BINS + t], (tile == 0 ? RS_FLAG_PRE : RS_FLAG_AGG) | cnt_d);
        if (EARLY && tile > 0) excl = lookback(cnt_d);
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
    if (!EARLY && t < RS_BINS) {
        if (tile > 0) excl = lookback(cnt_d);
        s_gbase[t] = offs[t] + excl - local_d;
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
            kout[dst] = kk;
            if constexpr (HAS_VALS) vout[dst] = s_vals[j];
        }
    }

}

// ---- persistent onesweep pass ----
// gridDim.x blocks (as many as are co-resident) loop over tiles taken in
// order from tile_ctr.  Right after a tile has been reordered into LDS, the
// block takes its next tile and loads its keys and payloads into the same
// registers, so that load's latency overlaps this tile's look-back and
// write phase.  A block waits (look-back) only on tiles taken before its
// own, all of which are held by running blocks, so every wait ends.
// NEXT: also count the digit (src >> nshift) & nmask of every row into
// g_nhist, the next pass's histogram (complete when this pass ends), so
// only the first pass needs a histogram read of the keys.  The counters
// are spread over RS_NCOPY LDS copies by lane.
constexpr int RS_NCOPY = 4;
template <typename V, bool HAS_VALS, int BLOCK, bool HASH, bool NEXT, int LBW = 16, int ITEMS = RS_ITEMS>
__global__ __launch_bounds__(BLOCK, 4) void k_rs_onesweep_pf(
    const uint32_t* __restrict__ kin, const V* __restrict__ vin, uint32_t* __restrict__ kout, V* __restrict__ vout,
    uint32_t n, uint32_t shift, uint32_t mask, const uint32_t* __restrict__ offs, uint32_t* __restrict__ status,
    uint32_t* __restrict__ tile_ctr, uint32_t ntiles, uint32_t nshift, uint32_t nmask,
    uint32_t* __restrict__ g_nhist) {
    constexpr int WAVES = BLOCK / PD_WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(TILE * 4 >= WAVES * RS_BINS * 8, "match masks must fit in the key buffer");
    __shared__ uint32_t s_next;
    __shared__ __align__(16) uint32_t s_keys[TILE];
    __shared__ __align__(16) V s_vals[HAS_VALS ? TILE : 1];
    __shared__ uint16_t s_cnt[WAVES][RS_BINS];
    __shared__ uint32_t s_off[RS_BINS];
    __shared__ uint32_t s_gbase[RS_BINS];
    __shared__ uint32_t s_wsum[RS_BINS / PD_WAVE];
    __shared__ uint32_t s_nh[NEXT ? RS_NCOPY : 1][NEXT ? RS_BINS : 1];
    auto s_match = reinterpret_cast<uint64_t (*)[RS_BINS]>(s_keys);

    const uint32_t t0 = threadIdx.x, l0 = pd::lane();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(t0 / PD_WAVE);
    if constexpr (NEXT)
        for (uint32_t i = t0; i < RS_NCOPY * RS_BINS; i += BLOCK) (&s_nh[0][0])[i] = 0;
    if (t0 == 0) s_next = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    uint32_t tile = s_next;

    uint32_t key[ITEMS];
    V val[HAS_VALS ? ITEMS : 1];
    uint32_t wrows = 0;   // rows of this wave in the current tile (wave-uniform)
    // wave-striped rows: wave_base + i*64 + lane, through buffer loads whose
    // range ends at the wave's last row (0 past it); rows past n get the
    // largest digit, so they rank after every real row and are never written
    // keys of a tile: loaded one tile ahead (during the previous tile's
    // look-back and writes); payloads: loaded when the tile's turn comes,
    // in flight during its ranking (so only the keys stay live across the
    // write phase)
    auto load_keys = [&](uint32_t tl, uint32_t lane4) {
        asm volatile("" : "+v"(lane4));   // lane * 4 + i * 256: immediate offsets, nothing hoisted
        const uint32_t wave_base = tl * (uint32_t)TILE + w * (uint32_t)(PD_WAVE * ITEMS);
        wrows = n > wave_base ? min((uint32_t)(PD_WAVE * ITEMS), n - wave_base) : 0u;
        const __amdgpu_buffer_rsrc_t rk = pd::buf_rsrc(kin + wave_base, wrows * 4u);
#pragma unroll
        for (int i = 0; i < ITEMS; i++)
            key[i] = __builtin_amdgcn_raw_buffer_load_b32(rk, lane4 + (uint32_t)i * PD_WAVE * 4u, 0, 0);
    };
    auto load_vals = [&](uint32_t tl, uint32_t lane4) {
        asm volatile("" : "+v"(lane4));
        const uint32_t wave_base = tl * (uint32_t)TILE + w * (uint32_t)(PD_WAVE * ITEMS);
        if constexpr (HAS_VALS) {
            const __amdgpu_buffer_rsrc_t rv = pd::buf_rsrc(vin + wave_base, wrows * (uint32_t)sizeof(V));
#pragma unroll
            for (int i = 0; i < ITEMS; i++) {
                if constexpr (sizeof(V) == 4)
                    val[i] = __builtin_amdgcn_raw_buffer_load_b32(rv, lane4 + (uint32_t)i * PD_WAVE * 4u, 0, 0);
                else
                    val[i] = __builtin_bit_cast(
                        V, __builtin_amdgcn_raw_buffer_load_b64(rv, 2u * lane4 + (uint32_t)i * PD_WAVE * 8u, 0, 0));
            }
        }
    };
    const __amdgpu_buffer_rsrc_t rko = pd::buf_rsrc(kout, n * 4u);
    const __amdgpu_buffer_rsrc_t rvo = pd::buf_rsrc(vout, HAS_VALS ? n * (uint32_t)sizeof(V) : 0u);
    if (tile < ntiles) load_keys(tile, l0 * 4u);

    while (tile < ntiles) {
        // lane id recomputed every tile by an opaque instruction pair, and
        // the thread id from it and the wave number: nothing derived from
        // them is hoisted out of the loop (and spilled; a scratch reload is
        // a vector-memory access, and waiting for it waits for every
        // earlier load and store of the wave)
        uint32_t l;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
        const uint32_t t = w * PD_WAVE + l;
        const uint64_t lt = (1ull << l) - 1ull;
        const uint32_t tile_count = min((uint32_t)TILE, n - tile * (uint32_t)TILE);
        // next ticket: issued now, its value is needed only after the
        // ranking.  The opaque zero offset keeps it a plain returning
        // atomic (a wave-aggregated one reads its result back at once).
        uint32_t my_next = 0;
        if (t == 0) {
            uint32_t z = 0;
            asm volatile("" : "+v"(z));
            my_next = atomicAdd(tile_ctr + z, 1u);
        }
        load_vals(tile, l * 4u);
        for (uint32_t i = l; i < RS_BINS; i += PD_WAVE) {
            s_match[w][i] = 0;
            s_cnt[w][i] = 0;
        }
        __syncthreads();
        auto digit_of = [&](int i) -> uint32_t {
            uint32_t kk = key[i];
            asm volatile("" : "+v"(kk));   // computed where used (see the single-tile pass)
            return ((uint32_t)i * PD_WAVE + l) < wrows ? ((rs_src<HASH>(kk) >> shift) & mask) : (RS_BINS - 1);
        };
        // ---- stable rank inside the wave (LDS match masks) ----
        uint32_t pos[ITEMS];
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
        if constexpr (NEXT) {
#pragma unroll
            for (int i = 0; i < ITEMS; i++)
                if (((uint32_t)i * PD_WAVE + l) < wrows)
                    atomicAdd(&s_nh[l % RS_NCOPY][(rs_src<HASH>(key[i]) >> nshift) & nmask], 1u);
        }
        __syncthreads();

        // ---- per digit: prefix over waves, tile count, publish the aggregate ----
        uint32_t cnt_d = 0, inc_d = 0;
        if (t < RS_BINS) {
#pragma unroll
            for (int ww = 0; ww < WAVES; ww++) {
                const uint32_t cw = s_cnt[ww][t];
                s_cnt[ww][t] = (uint16_t)cnt_d;
                cnt_d += cw;
            }
            pd::st_relaxed(&status[(uint64_t)tile * RS_BINS + t], (tile == 0 ? RS_FLAG_PRE : RS_FLAG_AGG) | cnt_d);
            inc_d = pd::wave_incl_scan(cnt_d);
            if (l == PD_WAVE - 1) s_wsum[w] = inc_d;
        }
        if (t == 0) s_next = my_next;
        __syncthreads();
        uint32_t local_d = 0;
        if (t < RS_BINS) {
            uint32_t add = 0;
            for (uint32_t i = 0; i < w; i++) add += s_wsum[i];
            local_d = add + inc_d - cnt_d;
            s_off[t] = local_d;
        }
        __syncthreads();

        // ---- reorder the tile by digit in LDS ----
#pragma unroll
        for (int i = 0; i < ITEMS; i++) {
            const uint32_t d = pos[i] >> 16;
            const uint32_t p = (pos[i] & 0xFFFFu) + s_off[d] + s_cnt[w][d];
            s_keys[p] = key[i];
            if constexpr (HAS_VALS) s_vals[p] = val[i];
        }
        // ---- first look-back window (thread t = digit t), then the next
        // tile's loads: the window is issued first, so waiting for it does
        // not wait for the prefetch ----
        const bool lb = t < RS_BINS && tile > 0;
        uint32_t st[LBW];
        if (lb) {
#pragma unroll
            for (int q = 0; q < LBW; q++)
                st[q] = ((int64_t)tile - 1 - q >= 0) ? pd::ld_relaxed(&status[(uint64_t)(tile - 1 - q) * RS_BINS + t])
                                                     : RS_FLAG_PRE;
        }
        PD_COMPILER_BARRIER();
        const uint32_t next = s_next;
        if (next < ntiles) load_keys(next, l * 4u);
        PD_COMPILER_BARRIER();

        // ---- per-digit look-back over earlier tiles ----
        if (t < RS_BINS) {
            uint32_t excl = 0;
            if (lb) {
                int64_t j = (int64_t)tile - 1;
                bool done = false, first = true;
                while (!done) {
                    if (!first) {
#pragma unroll
                        for (int q = 0; q < LBW; q++)
                            st[q] = (j - q >= 0) ? pd::ld_relaxed(&status[(uint64_t)(j - q) * RS_BINS + t]) : RS_FLAG_PRE;
                    }
                    first = false;
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
                pd::st_relaxed(&status[(uint64_t)tile * RS_BINS + t], RS_FLAG_PRE | (excl + cnt_d));
            }
            s_gbase[t] = offs[t] + excl - local_d;
        }
        __syncthreads();

        // ---- write keys and payloads in digit runs ----
#pragma unroll 4
        for (int k = 0; k < ITEMS; k++) {
            const uint32_t j = t + k * BLOCK;
            if (j >= tile_count) break;
            const uint32_t kk = s_keys[j];
            const uint32_t dst = s_gbase[(rs_src<HASH>(kk) >> shift) & mask] + j;
            __builtin_amdgcn_raw_buffer_store_b32(kk, rko, dst * 4u, 0, 0);
            if constexpr (HAS_VALS) {
                if constexpr (sizeof(V) == 4)
                    __builtin_amdgcn_raw_buffer_store_b32((uint32_t)s_vals[j], rvo, dst * 4u, 0, 0);
                else
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) uint32_t, s_vals[j]),
                                                          rvo, dst * 8u, 0, 0);
            }
        }
        __syncthreads();   // s_keys, s_vals, s_gbase are rewritten by the next tile
        tile = next;
    }
    if constexpr (NEXT) {
        for (uint32_t i = t0; i < RS_BINS; i += BLOCK) {
            uint32_t v = 0;
#pragma unroll
            for (int cp = 0; cp < RS_NCOPY; cp++) v += s_nh[cp][i];
            if (v) atomicAdd(&g_nhist[i], v);
        }
    }
}

// Look-back window of the passes (knob sort_lbw: 4, 8, 16 or 32).
static int sort_lbw(pimdal_ctx* c) {
    const int64_t v = c->tune_get("sort_lbw", 4);
    return v >= 32 ? 32 : v >= 16 ? 16 : v >= 8 ? 8 : 4;
}

// Co-resident blocks per CU of a kernel at a block size (at least 1).
template <typename K>
static uint32_t blocks_per_cu(K kernel, int block) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess || nb < 1) nb = 1;
    return (uint32_t)nb;
}

// Persistent passes.  hist_fused: only pass 0's histogram is in hist; each
// pass counts the next pass's digit and its offsets are scanned before it
// runs.
template <typename V, bool HV, bool HASH>
static void run_passes_pf(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                          const RsPassDesc& pdsc, uint32_t* hist, uint32_t* offs, uint32_t* ctrs, bool hist_fused,
                          hipStream_t s) {
    constexpr int B = 512;
    const int lbw = sort_lbw(c);
    const uint32_t ntiles = (uint32_t)ceil_div(n, (uint64_t)B * RS_ITEMS);
    const uint32_t P = pdsc.npass;
    Scratch tk(c, P > 1 ? n * 4 : 0, s);
    Scratch tv(c, (HV && P > 1) ? n * sizeof(V) : 0, s);
    Scratch st(c, (uint64_t)ntiles * RS_BINS * 4, s);
    const uint32_t* ksrc = kin;
    const V* vsrc = vin;
    for (uint32_t j = 0; j < P; j++) {
        const bool to_out = ((P - 1 - j) % 2) == 0;
        uint32_t* kdst = to_out ? kout : tk.as<uint32_t>();
        V* vdst = to_out ? vout : tv.as<V>();
        fill_u32(c, st.get(), 0, (uint64_t)ntiles * RS_BINS, s);
        if (hist_fused && j > 0)
            launch(c, "radix_scan", k_rs_scan, dim3(1), dim3(RS_BINS), 0, s, (const uint32_t*)(hist + j * RS_BINS),
                   offs + j * RS_BINS);
        const bool next = hist_fused && j + 1 < P;
        const uint32_t nshift = next ? pdsc.shift[j + 1] : 0, nmask = next ? pdsc.mask[j + 1] : 0;
        uint32_t* nh = next ? hist + (j + 1) * RS_BINS : nullptr;
        const char* name = HASH ? "part_onesweep" : "radix_onesweep";
#define PD_PF_LAUNCH(NX, W)                                                                                          \
    do {                                                                                                              \
        auto k = k_rs_onesweep_pf<V, HV, B, HASH, NX, W>;                                                             \
        const uint32_t g = std::min<uint32_t>(ntiles, (uint32_t)c->cu_count * blocks_per_cu(k, B));                   \
        launch(c, name, k, dim3(g), dim3(B), 0, s, ksrc, vsrc, kdst, vdst, (uint32_t)n, pdsc.shift[j], pdsc.mask[j],   \
               (const uint32_t*)(offs + j * RS_BINS), st.as<uint32_t>(), ctrs + j, ntiles, nshift, nmask, nh);         \
    } while (0)
        if (next) {
            if (lbw >= 16) PD_PF_LAUNCH(true, 16); else PD_PF_LAUNCH(true, 4);
        } else {
            if (lbw >= 16) PD_PF_LAUNCH(false, 16); else PD_PF_LAUNCH(false, 4);
        }
#undef PD_PF_LAUNCH
        ksrc = kdst;
        vsrc = vdst;
    }
}

template <typename V, bool HV, bool HASH = false>
static void run_passes(pimdal_ctx* c, const uint32_t* kin, const V* vin, uint64_t n, uint32_t* kout, V* vout,
                       const RsPassDesc& pdsc, const uint32_t* offs, uint32_t* ctrs, hipStream_t s) {
    const int64_t bk = c->tune_get("sort_block", 512);
    const int block = bk >= 1024 ? 1024 : bk <= 256 ? 256 : 512;
    // keys per thread: 16, or 8 (1024-thread blocks always take 8, so a
    // tile stays 8192 keys and two blocks still fit a CU's LDS)
    const int items = (block == 1024 || c->tune_get("sort_items", 16) == 8) ? 8 : 16;
    const bool early = c->tune_get("sort_early_lookback", 0) != 0;
    const bool ballot = c->tune_get("sort_rank", 0) == 1;
    const int lbw = sort_lbw(c);
    const uint64_t ntiles = ceil_div(n, (uint64_t)block * items);
    const uint32_t P = pdsc.npass;
    Scratch tk(c, P > 1 ? n * 4 : 0, s);
    Scratch tv(c, (HV && P > 1) ? n * sizeof(V) : 0, s);
    Scratch st(c, ntiles * RS_BINS * 4, s);
    const uint32_t* ksrc = kin;
    const V* vsrc = vin;
    const char* name = HASH ? "part_onesweep" : "radix_onesweep";
    for (uint32_t j = 0; j < P; j++) {
        const bool to_out = ((P - 1 - j) % 2) == 0;
        uint32_t* kdst = to_out ? kout : tk.as<uint32_t>();
        V* vdst = to_out ? vout : tv.as<V>();
        fill_u32(c, st.get(), 0, ntiles * RS_BINS, s);
#define PD_RS_LAUNCH(...)                                                                                         \
    launch(c, name, k_rs_onesweep<__VA_ARGS__>, dim3((uint32_t)ntiles), dim3(block), 0, s, ksrc, vsrc, kdst, vdst, n, \
           pdsc.shift[j], pdsc.mask[j], offs + j * RS_BINS, st.as<uint32_t>(), ctrs + j)
        if (block == 1024) PD_RS_LAUNCH(V, HV, 1024, false, HASH, false, 4, 8);
        else if (items == 8 && block == 512) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 4, 8);
        else if (block == 512 && !early && !ballot) {
            if (lbw == 32) PD_RS_LAUNCH(V, HV, 512, false, HASH, false, 32);
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

    // small scratch: hist[4*256] | offs[4*256] | andor[2] | ctr[4]
    Scratch small(c, (4 * RS_BINS * 2 + 8) * 4, s);
    uint32_t* hist = small.as<uint32_t>();
    uint32_t* offs = hist + 4 * RS_BINS;
    uint32_t* andor = offs + 4 * RS_BINS;
    uint32_t* ctrs = andor + 2;
    fill_u32(c, hist, 0, 4 * RS_BINS * 2, s);
    fill_u32(c, andor, 0xFFFFFFFFu, 1, s);
    fill_u32(c, andor + 1, 0, 1 + 4, s);

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
    const uint32_t hist_grid = (uint32_t)std::min<uint64_t>(ceil_div(n, 512 * 16), (uint64_t)c->cu_count * 4);
    // persistent passes; fused: the histogram read counts only the first
    // digit, and each pass counts the next one
    const bool persistent = c->tune_get("sort_persistent", 0) != 0;
    const bool fused = persistent && c->tune_get("sort_fused_hist", 1) != 0;
    RsPassDesc first{};
    first.npass = all.npass ? 1 : 0;
    first.shift[0] = all.shift[0];
    first.mask[0] = all.mask[0];
    if (all.npass)
        launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, kin, n,
               fused ? first : all, hist, andor);

    RsPassDesc run = all;
    if (detect && fused) {
        uint32_t ao[2];
        PD_HIP(hipMemcpyAsync(ao, andor, 8, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        const uint32_t varying = ao[0] ^ ao[1];
        run.npass = 0;
        for (int p = 0; p < 4; p++) {
            if ((varying >> (8 * p)) & 0xFF) {
                run.shift[run.npass] = 8 * p;
                run.mask[run.npass] = 0xFF;
                run.npass++;
            }
        }
        if (run.npass && run.shift[0] != 0) {
            // the lowest digit never varies: count the first varying one
            first.npass = 1;
            first.shift[0] = run.shift[0];
            first.mask[0] = 0xFF;
            fill_u32(c, hist, 0, RS_BINS, s);
            fill_u32(c, andor, 0xFFFFFFFFu, 1, s);
            launch(c, "radix_hist", k_rs_hist<false>, dim3(std::max(1u, hist_grid)), dim3(512), 0, s, kin, n, first,
                   hist, andor);
        }
    } else if (detect) {
        uint32_t ao[2];
        PD_HIP(hipMemcpyAsync(ao, andor, 8, hipMemcpyDeviceToHost, s));
        PD_HIP(hipStreamSynchronize(s));
        const uint32_t varying = ao[0] ^ ao[1];
        // keep only digits with a varying bit; move their histograms down
        run.npass = 0;
        for (int p = 0; p < 4; p++) {
            if ((varying >> (8 * p)) & 0xFF) {
                if (run.npass != (uint32_t)p)
                    copy_d2d(c, hist + run.npass * RS_BINS, hist + p * RS_BINS, RS_BINS * 4, s);
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
    launch(c, "radix_scan", k_rs_scan, dim3(fused ? 1 : run.npass), dim3(RS_BINS), 0, s, (const uint32_t*)hist, offs);
    if (persistent) {
        if (vw == 0)
            run_passes_pf<uint32_t, false, false>(c, kin, nullptr, n, kout, nullptr, run, hist, offs, ctrs, fused, s);
        else if (vw == 4)
            run_passes_pf<uint32_t, true, false>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, hist, offs,
                                                 ctrs, fused, s);
        else
            run_passes_pf<uint64_t, true, false>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, hist, offs,
                                                 ctrs, fused, s);
        return;
    }
    if (vw == 0)
        run_passes<uint32_t, false>(c, kin, nullptr, n, kout, nullptr, run, offs, ctrs, s);
    else if (vw == 4)
        run_passes<uint32_t, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, offs, ctrs, s);
    else
        run_passes<uint64_t, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, offs, ctrs, s);
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
    for (uint64_t v = a0 / 4 + t; v < v1; v += blockDim.x) {
        const uint4 q = reinterpret_cast<const uint4*>(keys)[v];
        atomicAdd(&h[pd::ref_hash(q.x) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.y) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.z) & (F - 1)], 1u);
        atomicAdd(&h[pd::ref_hash(q.w) & (F - 1)], 1u);
    }
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
    Scratch small(c, (4 * RS_BINS * 2 + 8) * 4, s);
    uint32_t* hist = small.as<uint32_t>();
    uint32_t* offs = hist + 4 * RS_BINS;
    uint32_t* andor = offs + 4 * RS_BINS;
    uint32_t* ctrs = andor + 2;
    fill_u32(c, hist, 0, 4 * RS_BINS * 2 + 8, s);
    // passes low digit first, widths as equal as possible (longer write runs)
    RsPassDesc run{};
    run.npass = (bits + 7) / 8;
    for (uint32_t p = 0, b = 0; p < run.npass; p++) {
        const uint32_t w = (bits - b) / (run.npass - p);
        run.shift[p] = b;
        run.mask[p] = (1u << w) - 1;
        b += w;
    }
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
    launch(c, "radix_scan", k_rs_scan, dim3(run.npass), dim3(RS_BINS), 0, s, (const uint32_t*)hist, offs);
    if (c->tune_get("sort_persistent", 0) != 0) {
        if (vw == 4)
            run_passes_pf<uint32_t, true, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, hist, offs,
                                                ctrs, false, s);
        else if (vw == 8)
            run_passes_pf<uint64_t, true, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, hist, offs,
                                                ctrs, false, s);
        else
            run_passes_pf<uint32_t, false, true>(c, kin, nullptr, n, kout, nullptr, run, hist, offs, ctrs, false, s);
    } else if (vw == 4)
        run_passes<uint32_t, true, true>(c, kin, (const uint32_t*)vin, n, kout, (uint32_t*)vout, run, offs, ctrs, s);
    else if (vw == 8)
        run_passes<uint64_t, true, true>(c, kin, (const uint64_t*)vin, n, kout, (uint64_t*)vout, run, offs, ctrs, s);
    else
        run_passes<uint32_t, false, true>(c, kin, nullptr, n, kout, nullptr, run, offs, ctrs, s);
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
