// gen.hip — device side of the counter-based synthetic data generators
// (include/pimdal_gen.h); the host restatement produces identical values.
#include <vector>
#include <cmath>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "../../include/pimdal_gen.h"

namespace pd {

#define PD_GRID_STRIDE(i, n) \
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_gen_u32(uint64_t st, uint64_t n, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = (uint32_t)(pd_rand(st, i) >> 32);
}

template <typename T>
__global__ void k_gen_range(uint64_t st, uint64_t n, int64_t lo, uint64_t span, T* out) {
    PD_GRID_STRIDE(i, n) out[i] = (T)(lo + (int64_t)pd_uniform_span(pd_rand(st, i), span));
}

__global__ void k_gen_perm(uint64_t st, uint64_t m, uint64_t off, uint64_t n, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = pd_perm(st, m, off + i);
}

__global__ void k_perm_apply(uint64_t st, uint64_t m, uint64_t n, int64_t delta, const uint32_t* in, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = pd_perm(st, m, (uint64_t)((int64_t)in[i] + delta));
}

__global__ void k_iota(uint64_t n, uint32_t start, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = start + (uint32_t)i;
}

__global__ void k_gen_zipf(uint64_t st, const double* cdf, uint64_t m, uint64_t n, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = pd_zipf_rank(cdf, m, pd_rand(st, i));
}

__global__ void k_fmix32(uint64_t n, const uint32_t* in, uint32_t* out) {
    PD_GRID_STRIDE(i, n) out[i] = pd_fmix32(in[i]);
}

__global__ void k_gen_orderkey(uint64_t sr, uint64_t sv, uint64_t n, uint64_t groups, int32_t* out) {
    PD_GRID_STRIDE(i, n) {
        const uint64_t g = (uint64_t)pd_perm(sr, n, i) % groups;
        out[i] = (int32_t)pd_perm(sv, 1ull << 31, g);
    }
}

static dim3 gen_grid(pimdal_ctx* c, uint64_t n) {
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), (uint64_t)c->cu_count * 16)));
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_gen_u32(pimdal_ctx* c, uint64_t seed, uint32_t col, uint64_t n, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    launch(c, "gen", k_gen_u32, gen_grid(c, n), dim3(256), 0, c->stream, pd_stream(seed, col), n, out);
    PD_API_END
}

extern "C" int pimdal_gen_range(pimdal_ctx* c, pimdal_type type, uint64_t seed, uint32_t col, uint64_t n, int64_t lo,
                                int64_t hi, void* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(hi >= lo && (uint64_t)(hi - lo) < (1ull << 32), PIMDAL_E_INVALID, "range must satisfy 0 <= hi-lo < 2^32");
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    const uint64_t st = pd_stream(seed, col);
    switch (type) {
        case PIMDAL_U8: launch(c, "gen", k_gen_range<uint8_t>, gen_grid(c, n), dim3(256), 0, c->stream, st, n, lo, span, (uint8_t*)out); break;
        case PIMDAL_I32: launch(c, "gen", k_gen_range<int32_t>, gen_grid(c, n), dim3(256), 0, c->stream, st, n, lo, span, (int32_t*)out); break;
        case PIMDAL_U32: launch(c, "gen", k_gen_range<uint32_t>, gen_grid(c, n), dim3(256), 0, c->stream, st, n, lo, span, (uint32_t*)out); break;
        case PIMDAL_I64: launch(c, "gen", k_gen_range<int64_t>, gen_grid(c, n), dim3(256), 0, c->stream, st, n, lo, span, (int64_t*)out); break;
        case PIMDAL_U64: launch(c, "gen", k_gen_range<uint64_t>, gen_grid(c, n), dim3(256), 0, c->stream, st, n, lo, span, (uint64_t*)out); break;
        default: fail(PIMDAL_E_UNSUPPORTED, "unsupported generator type");
    }
    PD_API_END
}

extern "C" int pimdal_gen_perm(pimdal_ctx* c, uint64_t seed, uint32_t col, uint64_t m, uint64_t off, uint64_t n, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(m >= 1 && m <= (1ull << 32), PIMDAL_E_INVALID, "permutation domain must be in [1, 2^32]");
    launch(c, "gen", k_gen_perm, gen_grid(c, n), dim3(256), 0, c->stream, pd_stream(seed, col), m, off, n, out);
    PD_API_END
}

extern "C" int pimdal_perm_apply(pimdal_ctx* c, uint64_t seed, uint32_t col, uint64_t m, uint64_t n, int64_t delta,
                                 const uint32_t* in, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && ((in && out) || !n), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(m >= 1 && m <= (1ull << 32), PIMDAL_E_INVALID, "permutation domain must be in [1, 2^32]");
    launch(c, "gen", k_perm_apply, gen_grid(c, n), dim3(256), 0, c->stream, pd_stream(seed, col), m, n, delta, in, out);
    PD_API_END
}

extern "C" int pimdal_gen_iota(pimdal_ctx* c, uint64_t n, uint32_t start, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    launch(c, "gen", k_iota, gen_grid(c, n), dim3(256), 0, c->stream, n, start, out);
    PD_API_END
}

extern "C" int pimdal_gen_zipf(pimdal_ctx* c, uint64_t seed, uint32_t col, uint64_t m, double s, uint64_t n, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(m >= 1 && m < (1ull << 32), PIMDAL_E_INVALID, "zipf domain must be in [1, 2^32)");
    // cdf[k-1] = sum_{j<=k} j^-s, summed left to right in double (same loop as the host restatement)
    std::vector<double> cdf(m);
    double acc = 0.0;
    for (uint64_t k = 1; k <= m; k++) {
        acc += 1.0 / std::pow((double)k, s);
        cdf[k - 1] = acc;
    }
    Scratch d(c, m * 8);
    PD_HIP(hipMemcpyAsync(d.get(), cdf.data(), m * 8, hipMemcpyHostToDevice, c->stream));
    launch(c, "gen", k_gen_zipf, gen_grid(c, n), dim3(256), 0, c->stream, pd_stream(seed, col), d.as<const double>(), m, n, out);
    PD_HIP(hipStreamSynchronize(c->stream));
    PD_API_END
}

extern "C" int pimdal_fmix32(pimdal_ctx* c, uint64_t n, const uint32_t* in, uint32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && ((in && out) || !n), PIMDAL_E_INVALID, "NULL argument");
    launch(c, "gen", k_fmix32, gen_grid(c, n), dim3(256), 0, c->stream, n, in, out);
    PD_API_END
}

extern "C" int pimdal_gen_orderkey(pimdal_ctx* c, uint64_t seed, uint64_t n, uint64_t groups, int32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !n), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(groups >= 1 && n <= (1ull << 32), PIMDAL_E_INVALID, "bad orderkey shape");
    launch(c, "gen", k_gen_orderkey, gen_grid(c, n), dim3(256), 0, c->stream, pd_stream(seed, PD_COL_ORDERKEY_ROWS),
           pd_stream(seed, PD_COL_ORDERKEY_VALS), n, groups, out);
    PD_API_END
}

namespace pd {
__global__ void k_gen_dim_keys(uint64_t sd, uint64_t sv, uint64_t groups, int32_t* out) {
    PD_GRID_STRIDE(i, groups) out[i] = (int32_t)pd_perm(sv, 1ull << 31, pd_perm(sd, groups, i));
}
}  // namespace pd

extern "C" int pimdal_gen_dim_keys(pimdal_ctx* c, uint64_t seed, uint64_t groups, int32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c && (out || !groups), PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(groups >= 1 && groups <= (1ull << 31), PIMDAL_E_INVALID, "bad group count");
    launch(c, "gen", k_gen_dim_keys, gen_grid(c, groups), dim3(256), 0, c->stream, pd_stream(seed, PD_COL_DIM_PERM),
           pd_stream(seed, PD_COL_ORDERKEY_VALS), groups, out);
    PD_API_END
}
