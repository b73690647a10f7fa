// autotune.hip — picks tile shapes, look-back order, partition fan-out and
// chunk sizes per operator by timing candidates on synthetic data of the
// caller's size (device-generated, so nothing crosses PCIe), keeps the
// fastest in the context's tuning table and appends it to a cache file
// ("op n key value" lines) that pimdal_ctx_load_tuning reads back.
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>

#include "pd_runtime.hpp"
#include "pd_ops.hpp"
#include "../../include/pimdal_gen.h"

namespace pd {

namespace {

struct Knob {
    std::string key;
    std::vector<int64_t> values;
};

// Times `run` (already warmed up) with hipEvents on the context stream.
template <typename F>
double time_ms(pimdal_ctx* c, F&& run, int reps) {
    run();
    PD_HIP(hipStreamSynchronize(c->stream));
    hipEvent_t e0, e1;
    PD_HIP(hipEventCreate(&e0));
    PD_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int r = 0; r < reps; r++) {
        PD_HIP(hipEventRecord(e0, c->stream));
        run();
        PD_HIP(hipEventRecord(e1, c->stream));
        PD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PD_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

__global__ void k_tune_fill(uint32_t* a, uint64_t n, uint64_t seed, uint32_t mod) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(pd_rand(seed, i) >> 32);
        a[i] = mod ? r % mod : r;
    }
}

void fill(pimdal_ctx* c, void* a, uint64_t n, uint64_t seed, uint32_t mod) {
    launch(c, "tune_fill", k_tune_fill, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), 8192))),
           dim3(256), 0, c->stream, (uint32_t*)a, n, seed, mod);
}

// Coordinate descent over the knobs: each knob in turn, the others fixed.
template <typename F>
void tune(pimdal_ctx* c, const std::vector<Knob>& knobs, F&& run, int reps) {
    for (const Knob& k : knobs) {
        double best = 1e30;
        int64_t arg = k.values.front();
        for (int64_t v : k.values) {
            c->tune[k.key] = v;
            const double ms = time_ms(c, run, reps);
            if (ms < best) { best = ms; arg = v; }
        }
        c->tune[k.key] = arg;
    }
}

}  // namespace

}  // namespace pd

using namespace pd;

extern "C" int pimdal_autotune(pimdal_ctx* c, const char* op, uint64_t n, const char* cache_path) {
    PD_API_BEGIN
    PD_REQUIRE(c && op && n > 0, PIMDAL_E_INVALID, "bad argument");
    PD_NO_CAPTURE(c);
    hipStream_t s = c->stream;
    std::vector<Knob> knobs;
    const std::string o(op);
    if (o == "sort") {
        PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "sort tuning size too large");
        Scratch k(c, n * 4), v(c, n * 4), ko(c, n * 4), vo(c, n * 4);
        fill(c, k.get(), n, 1, 0);
        fill(c, v.get(), n, 2, 0);
        knobs = {{"sort_xcd_tiles", {16, 8, 32, 0}}, {"sort_block", {512, 256}}, {"sort_early_lookback", {0, 1}}};
        tune(c, knobs, [&] { sort_pairs(c, k.as<uint32_t>(), v.get(), 4, n, ko.as<uint32_t>(), vo.get(), 0, 32, s); }, 3);
    } else if (o == "select") {
        Scratch k(c, n * 4), p(c, n * 4), op_(c, n * 4), orow(c, n * 4), cnt(c, 8);
        fill(c, k.get(), n, 3, 0);
        fill(c, p.get(), n, 4, 0);
        knobs = {{"select_mode", {0, 1}}, {"select_block", {256, 512, 1024}}, {"select_lookback_width", {1, 4}}};
        tune(c, knobs, [&] {
            select_range(c, PIMDAL_U32, k.get(), p.as<uint32_t>(), n, 0, 429496730, op_.as<uint32_t>(),
                         orow.as<uint32_t>(), cnt.as<uint64_t>(), s);
        }, 5);
    } else if (o == "hash_join") {
        const uint64_t nb = std::max<uint64_t>(1, n / 8);
        Scratch bk(c, nb * 4), bp(c, nb * 4), pk(c, n * 4), pp(c, n * 4), o1(c, n * 4), o2(c, n * 4);
        fill(c, bk.get(), nb, 5, 0);
        fill(c, bp.get(), nb, 6, 0);
        fill(c, pk.get(), n, 7, 0);
        fill(c, pp.get(), n, 8, 0);
        // 0 / -1 stand for the size-aware defaults, so a cached value keeps
        // adapting to the size of later calls; the one global table (hj_mode
        // 1) is not a candidate: above 2^16 build rows its probe reads whole
        // lines from HBM per key and loses to the partitioned join
        knobs = {{"hj_mode", {-1, 0}}, {"hj_block", {512, 1024}}, {"hj_rows_per_part", {0, 4096}},
                 {"hj_chunk", {0, 16384, 32768, 65536}}};
        tune(c, knobs, [&] {
            hash_join(c, bk.as<uint32_t>(), bp.as<uint32_t>(), nb, pk.as<uint32_t>(), pp.as<uint32_t>(), n,
                      o1.as<uint32_t>(), o2.as<uint32_t>(), n, s);
        }, 2);
    } else if (o == "groupby") {
        Scratch k(c, n * 4), v(c, n * 8), ok(c, n * 4), os(c, n * 8), oc(c, n * 8);
        fill(c, k.get(), n, 9, (uint32_t)std::max<uint64_t>(1, n / 4));
        fill(c, v.get(), 2 * n, 10, 1000);
        knobs = {{"gb_rows_per_part", {2048, 4096, 8192}}};
        const int64_t hint = c->tune_get("gb_groups_hint", 0);
        c->tune["gb_groups_hint"] = (int64_t)(n / 4);
        tune(c, knobs, [&] {
            hash_aggr(c, k.as<uint32_t>(), v.as<int64_t>(), n, ok.as<uint32_t>(), os.as<int64_t>(), oc.as<int64_t>(), n, s);
        }, 2);
        c->tune["gb_groups_hint"] = hint;
    } else if (o == "q1") {
        // the filter + group-by by orderkey over a fact table of n rows (four
        // rows per key, measures in their TPC-H ranges): partition fan-out
        // (rows per partition) and LDS table size of the per-partition tables
        PD_REQUIRE(n < (1ull << 30), PIMDAL_E_INVALID, "q1 tuning size too large");
        Scratch ok(c, n * 4), qt(c, n * 8), pr(c, n * 8), di(c, n * 8), tx(c, n * 8), dt(c, n * 4);
        const uint64_t seed = 0x5EED0004ull;
        auto chk = [](int rc) { PD_REQUIRE(rc == PIMDAL_OK, rc, pimdal_last_error()); };
        chk(pimdal_gen_orderkey(c, seed, n, std::max<uint64_t>(1, n / 4), ok.as<int32_t>()));
        chk(pimdal_gen_range(c, PIMDAL_I64, seed, 12, n, 1, 50, qt.get()));
        chk(pimdal_gen_range(c, PIMDAL_I64, seed, 13, n, 90000, 10500000, pr.get()));
        chk(pimdal_gen_range(c, PIMDAL_I64, seed, 14, n, 0, 10, di.get()));
        chk(pimdal_gen_range(c, PIMDAL_I64, seed, 15, n, 0, 8, tx.get()));
        chk(pimdal_gen_range(c, PIMDAL_I32, seed, 16, n, 0, 2556, dt.get()));
        Scratch rk(c, n * 4), r1(c, n * 8), r2(c, n * 8), r3(c, n * 8), r4(c, n * 8), r5(c, n * 16), r6(c, n * 8);
        const pimdal_q1_result res{rk.as<int32_t>(), r1.as<int64_t>(), r2.as<int64_t>(), r3.as<int64_t>(),
                                   r4.as<int64_t>(), r5.as<int64_t>(), r6.as<int64_t>()};
        knobs = {{"gb_rows_per_part", {2048, 1024, 4096}}, {"gb_q1_slots", {1024, 2048}}};
        tune(c, knobs, [&] {
            filter_groupby_q1(c, dt.as<int32_t>(), nullptr, nullptr, ok.as<int32_t>(), qt.as<int64_t>(),
                              pr.as<int64_t>(), di.as<int64_t>(), tx.as<int64_t>(), n, 2505, res, n, s);
        }, 2);
    } else {
        fail(PIMDAL_E_INVALID, "unknown operator '" + o + "' (sort, select, hash_join, groupby, q1)");
    }
    if (cache_path && *cache_path) {
        FILE* f = fopen(cache_path, "a");
        PD_REQUIRE(f, PIMDAL_E_INVALID, std::string("cannot open tuning cache ") + cache_path);
        for (const Knob& k : knobs) fprintf(f, "%s %llu %s %lld\n", op, (unsigned long long)n, k.key.c_str(), (long long)c->tune[k.key]);
        fclose(f);
    }
    PD_API_END
}

extern "C" int pimdal_ctx_load_tuning(pimdal_ctx* c, const char* cache_path) {
    PD_API_BEGIN
    PD_REQUIRE(c && cache_path, PIMDAL_E_INVALID, "NULL argument");
    FILE* f = fopen(cache_path, "r");
    PD_REQUIRE(f, PIMDAL_E_INVALID, std::string("cannot open tuning cache ") + cache_path);
    char op[64], key[128];
    unsigned long long n;
    long long v;
    while (fscanf(f, "%63s %llu %127s %lld", op, &n, key, &v) == 4) c->tune[key] = v;  // later lines win
    fclose(f);
    PD_API_END
}
