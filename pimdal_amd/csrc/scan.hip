// scan.hip — device-wide exclusive prefix sum of uint32 counts in one pass
// (decoupled look-back over 4096-element tiles).  Used for partition
// offsets, work-item offsets and output offsets.  The last element of the
// output array pair is the total: out[n] receives the sum when out has room
// for n + 1 entries (always the case for callers in this library).
#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr int SC_BLOCK = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;

__global__ __launch_bounds__(SC_BLOCK) void k_scan_excl(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        uint64_t n, uint64_t* __restrict__ status,
                                                        uint32_t* __restrict__ tile_ctr, uint64_t ntiles,
                                                        uint64_t* __restrict__ total) {
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_wsum[SC_BLOCK / PD_WAVE];
    __shared__ uint64_t s_excl;
    const uint32_t t = threadIdx.x, w = t / PD_WAVE, l = pd::lane();
    if (t == 0) s_tile = atomicAdd(tile_ctr, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t base = tile * SC_TILE + (uint64_t)t * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint64_t sum = 0;
    // 16-byte accesses where the thread's 16 elements are whole and aligned
    const bool vec = base + SC_ITEMS <= n && ((uintptr_t)(in + base) % 16) == 0 && ((uintptr_t)(out + base) % 16) == 0;
    if (vec) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int i = 0; i < SC_ITEMS / 4; i++) {
            const uint4 x = p[i];
            v[4 * i] = x.x;
            v[4 * i + 1] = x.y;
            v[4 * i + 2] = x.z;
            v[4 * i + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SC_ITEMS; i++) v[i] = (base + i < n) ? in[base + i] : 0u;
    }
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) sum += v[i];
    const uint64_t inc = pd::wave_incl_scan64(sum);
    if (l == PD_WAVE - 1) s_wsum[w] = inc;
    __syncthreads();
    uint64_t wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SC_BLOCK / PD_WAVE; i++) {
        wpre += (i < (int)w) ? s_wsum[i] : 0;
        tot += s_wsum[i];
    }
    if (w == 0) {
        const uint64_t e = pd::lookback_single<4>(status, tile, tot);
        if (l == 0) {
            s_excl = e;
            if (tile == ntiles - 1 && total) *total = e + tot;
        }
    }
    __syncthreads();
    uint64_t run = s_excl + wpre + inc - sum;
    uint32_t o[SC_ITEMS];
#pragma unroll
    for (int i = 0; i < SC_ITEMS; i++) {
        o[i] = (uint32_t)run;
        run += v[i];
    }
    if (vec) {
        uint4* q = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int i = 0; i < SC_ITEMS / 4; i++) q[i] = make_uint4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < SC_ITEMS; i++)
            if (base + i < n) out[base + i] = o[i];
    }
    if (tile == ntiles - 1 && t == SC_BLOCK - 1) out[n] = (uint32_t)(s_excl + tot);
}

// out has n + 1 entries; out[n] = total.  total_dev (optional) gets the
// 64-bit total.  in and out may alias (each element is read before the
// same thread writes it, and tiles only read their own range).
void exclusive_scan_u32(pimdal_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total_dev,
                        hipStream_t s) {
    if (n == 0) {
        fill_u32(c, out, 0, 1, s);
        if (total_dev) fill_u64(c, total_dev, 0, 1, s);
        return;
    }
    const uint64_t ntiles = ceil_div(n, SC_TILE);
    Scratch st(c, 8 * (ntiles + 1), s);
    fill_u64(c, st.get(), 0, ntiles + 1, s);
    launch(c, "scan", k_scan_excl, dim3((uint32_t)ntiles), dim3(SC_BLOCK), 0, s, in, out, n, st.as<uint64_t>(),
           reinterpret_cast<uint32_t*>(st.as<uint64_t>() + ntiles), ntiles, total_dev);
}

}  // namespace pd
