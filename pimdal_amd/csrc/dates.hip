// dates.hip — packed dates (pimdal/main/tpc_h/reader/read_table.cpp:8-16,
// date_to_int): day + (month << 5) + ((year - 1900) << 9).
//
// The TPC-H programs compare date columns and literals in this packed
// form (pimdal/main/tpc_h/query6/host/host_q6.cpp:46-47).  Literals come
// through pimdal_date_pack / pimdal_date_pack_str; Arrow date32 columns
// (int32 days since 1970-01-01) through k_date32_pack, one streaming pass
// (4 B read + 4 B written per value, HBM-bound).
//
// Days to (year, month, day): H. Hinnant, "chrono-Compatible Low-Level Date
// Algorithms", civil_from_days: shift the epoch to 0000-03-01 so the leap
// day ends each 400-year era's years, then split the day of the era into
// year of era, day of year and a March-based month by integer division.
#include <cctype>
#include <climits>

#include "pd_runtime.hpp"
#include "pd_device.hpp"

namespace pd {

// (year - 1900) << 9 as a multiplication, so years before 1900 (a negative
// tm_year) pack the same way as a two's-complement shift of tm_year; the
// sum is taken modulo 2^32, as 32-bit int arithmetic gives it for years too
// far from 1900 to fit.
__host__ __device__ __forceinline__ int32_t date_pack_ymd(int64_t y, int32_t m, int32_t d) {
    return (int32_t)((uint32_t)d + (uint32_t)m * 32u + (uint32_t)(uint64_t)(y - 1900) * 512u);
}

// Days since 0000-03-01 (z) to the packed date; Z is the type z and the
// era are computed in.
template <typename Z>
__host__ __device__ __forceinline__ int32_t civil_packed(Z z) {
    const Z era = (z >= 0 ? z : z - 146096) / 146097;             // 400-year eras, floor
    const uint32_t doe = (uint32_t)(z - era * 146097);            // [0, 146096]
    const uint32_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;   // [0, 399]
    const uint32_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100); // [0, 365], from March 1
    const uint32_t mp = (5 * doy + 2) / 153;                      // [0, 11], March = 0
    const int32_t d = (int32_t)(doy - (153 * mp + 2) / 5 + 1);    // [1, 31]
    const int32_t m = (int32_t)(mp < 10 ? mp + 3 : mp - 9);       // [1, 12]
    const int64_t y = (int64_t)yoe + (int64_t)era * 400 + (m <= 2 ? 1 : 0);
    return date_pack_ymd(y, m, d);
}

// Any int32 day count: in 32 bits while days + 719468 fits (every date
// before the year 5,881,000), else the day number and era in 64 bits.
__host__ __device__ __forceinline__ int32_t date32_to_packed(int32_t days) {
    if (days <= INT32_MAX - 719468) return civil_packed<int32_t>(days + 719468);
    return civil_packed<int64_t>((int64_t)days + 719468);
}

// Four values per thread per step with 16-byte loads and stores when both
// pointers are 16-byte aligned; in place is fine (each thread reads its
// values before writing the same positions).
template <bool VEC>
__global__ __launch_bounds__(256) void k_date32_pack(const int32_t* days, uint64_t n, int32_t* out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = VEC ? n / 4 : 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n4; v += stride) {
        const int4 x = reinterpret_cast<const int4*>(days)[v];
        int4 y;
        y.x = date32_to_packed(x.x);
        y.y = date32_to_packed(x.y);
        y.z = date32_to_packed(x.z);
        y.w = date32_to_packed(x.w);
        reinterpret_cast<int4*>(out)[v] = y;
    }
    for (uint64_t i = n4 * 4 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = date32_to_packed(days[i]);
}

// Up to `maxd` decimal digits (at least one) into *v; returns the position after them.
static const char* parse_digits(const char* p, int maxd, int32_t* v) {
    int k = 0;
    int32_t x = 0;
    while (k < maxd && *p >= '0' && *p <= '9') {
        x = x * 10 + (*p - '0');
        p++;
        k++;
    }
    if (k == 0) return nullptr;
    *v = x;
    return p;
}

}  // namespace pd

using namespace pd;

extern "C" int pimdal_date_pack(int32_t year, int32_t month, int32_t day, int32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(out, PIMDAL_E_INVALID, "NULL argument");
    PD_REQUIRE(month >= 1 && month <= 12, PIMDAL_E_INVALID, "month out of 1..12");
    PD_REQUIRE(day >= 1 && day <= 31, PIMDAL_E_INVALID, "day out of 1..31");
    PD_REQUIRE(year > 1900 - (1 << 22) && year < 1900 + (1 << 22), PIMDAL_E_INVALID,
               "year too far from 1900 for a 32-bit packed date");
    *out = date_pack_ymd(year, month, day);
    PD_API_END
}

extern "C" int pimdal_date_pack_str(const char* s, int32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(s && out, PIMDAL_E_INVALID, "NULL argument");
    const char* p = s;
    while (std::isspace((unsigned char)*p)) p++;
    int32_t y = 0, m = 0, d = 0;
    p = parse_digits(p, 4, &y);
    PD_REQUIRE(p && *p == '-', PIMDAL_E_INVALID, std::string("not a YYYY-MM-DD date: '") + s + "'");
    p = parse_digits(p + 1, 2, &m);
    PD_REQUIRE(p && *p == '-', PIMDAL_E_INVALID, std::string("not a YYYY-MM-DD date: '") + s + "'");
    p = parse_digits(p + 1, 2, &d);
    PD_REQUIRE(p, PIMDAL_E_INVALID, std::string("not a YYYY-MM-DD date: '") + s + "'");
    while (std::isspace((unsigned char)*p)) p++;
    PD_REQUIRE(*p == '\0', PIMDAL_E_INVALID, std::string("not a YYYY-MM-DD date: '") + s + "'");
    return pimdal_date_pack(y, m, d, out);
    PD_API_END
}

extern "C" int pimdal_date32_pack_dev(pimdal_ctx* c, const int32_t* days, uint64_t n, int32_t* out) {
    PD_API_BEGIN
    PD_REQUIRE(c, PIMDAL_E_INVALID, "NULL context");
    if (n == 0) return PIMDAL_OK;
    PD_REQUIRE(days && out, PIMDAL_E_INVALID, "NULL column");
    PD_REQUIRE(out == days || (const char*)out + n * 4 <= (const char*)days ||
                   (const char*)days + n * 4 <= (const char*)out,
               PIMDAL_E_INVALID, "input and output overlap without being the same column");
    const bool vec = (uintptr_t)days % 16 == 0 && (uintptr_t)out % 16 == 0;
    const uint64_t work = vec ? ceil_div(n, 4) : n;
    const dim3 grid((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(work, 256), (uint64_t)c->cu_count * 8)));
    if (vec)
        launch(c, "date32_pack", k_date32_pack<true>, grid, dim3(256), 0, c->stream, days, n, out);
    else
        launch(c, "date32_pack", k_date32_pack<false>, grid, dim3(256), 0, c->stream, days, n, out);
    PD_API_END
}
