// pd_ops.hpp — internal host entry points of the operators (stream-ordered,
// throwing pd::Error); the C ABI wraps these.
#pragma once
#include "pd_runtime.hpp"

namespace pd {

void select_range(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                  int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s);

// The one-pass select with its status words in a static device array
// instead of pool memory (no allocation); select_noalloc_init looks up that
// array's address on the current device, once per device.
void select_range_noalloc(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                          int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s);
void select_noalloc_init(int device);

void sort_pairs(pimdal_ctx* c, const uint32_t* kin, const void* vin, uint32_t vw, uint64_t n, uint32_t* kout,
                void* vout, int begin_bit, int end_bit, hipStream_t s);

void exclusive_scan_u32(pimdal_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total_dev,
                        hipStream_t s);

// Stable partition of (key, payload) rows by ref_hash(key) & (2^bits - 1)
// (onesweep passes over hash digits); payloads are vw = 0, 4 or 8 bytes;
// bounds[f] .. bounds[f + 1] is partition f, bounds has 2^bits + 1 entries.
void hash_partition(pimdal_ctx* c, const uint32_t* kin, const void* vin, uint32_t vw, uint64_t n, uint32_t bits,
                    uint32_t* kout, void* vout, uint32_t* bounds, hipStream_t s);

// One pass into 2^bits <= 1024 partitions by ref_hash(key) & (2^bits - 1)
// (partition.hip): (key, payload) rows out, order inside a partition
// unspecified; bounds as hash_partition's.
void hash_partition_1p(pimdal_ctx* c, const uint32_t* kin, const uint32_t* vin, uint64_t n, uint32_t bits, uint2* out,
                       uint32_t* bounds, hipStream_t s);

// Per-tile partition counts cnt[tile][1024] (columns below F used) to
// per-tile offsets in place, and start[p] (p <= F) = rows of partitions
// below p (partition.hip).
void partition_tile_scan(pimdal_ctx* c, uint32_t* cnt, uint32_t ntiles, uint32_t F, uint32_t* start, hipStream_t s);

// The last pass of hash_partition for rows already partitioned by the low
// lo_bits hash bits: one stable pass over bits [lo_bits, bits) from that
// digit's histogram (8-byte payloads), then the bounds (radix_sort.hip).
void hash_partition_finish(pimdal_ctx* c, const uint32_t* kin, const uint64_t* vin, uint64_t n, uint32_t bits,
                           uint32_t lo_bits, const uint32_t* hist_hi, uint32_t* kout, uint64_t* vout, uint32_t* bounds,
                           hipStream_t s);

uint64_t hash_join(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                   const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap, hipStream_t s);

uint64_t sort_merge_join(pimdal_ctx* c, const uint32_t* bk, const uint32_t* bp, uint64_t nb, const uint32_t* pk,
                         const uint32_t* pp, uint64_t np, uint32_t* out_p, uint32_t* out_b, uint64_t cap,
                         hipStream_t s);

uint64_t hash_aggr(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                   int64_t* out_sum, int64_t* out_cnt, uint64_t cap, hipStream_t s);

uint64_t sort_aggr(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                   int64_t* out_sum, int64_t* out_cnt, uint64_t cap, hipStream_t s);

// SUM and COUNT per uint32 key (aggr_sum.hip).  aggr_sum_small: the
// few-groups path, false (outputs untouched) when the groups do not fit its
// tables; ordered = groups in ascending key order.  aggr_sum_sorted: rows
// already sorted by key, groups in key order; both throw PIMDAL_E_CAPACITY
// past cap groups.
bool aggr_sum_small(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey, int64_t* osum,
                    int64_t* ocnt, uint64_t cap, bool ordered, uint64_t* groups, hipStream_t s);
uint64_t aggr_sum_sorted(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                         int64_t* osum, int64_t* ocnt, uint64_t cap, hipStream_t s);
// Keys within 2^24 of the smallest (knob aggr_dense_max): one range-partition
// pass and per-partition LDS arrays indexed by key; groups in key order.
// False (outputs untouched) when the key range is wider.
bool aggr_sum_dense(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                    int64_t* osum, int64_t* ocnt, uint64_t cap, uint64_t* groups, hipStream_t s);
// Keys within 8192 values of each other (by a sample of the first 2^16 rows;
// knob aggr_direct): per-block LDS arrays indexed by key, no table and no
// partition pass; groups in key order.  False (outputs untouched) when the
// sample or any key falls outside.
bool aggr_sum_direct(pimdal_ctx* c, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* okey,
                     int64_t* osum, int64_t* ocnt, uint64_t cap, uint64_t* groups, hipStream_t s);

uint64_t filter_groupby_q1(pimdal_ctx* c, const int32_t* date, const uint8_t* f1, const uint8_t* f2,
                           const int32_t* key, const int64_t* qty, const int64_t* price, const int64_t* disc,
                           const int64_t* tax, uint64_t n, int32_t cutoff, const pimdal_q1_result& out,
                           uint64_t cap, hipStream_t s);

// Dense codes of two byte columns: lut_dev[0..512) = code of each byte value
// (flag1 then flag2), lut_dev[512..1024) = inverse; meta = {n1, n2}.
void flag_codes(pimdal_ctx* c, const uint8_t* f1, const uint8_t* f2, uint64_t n, uint8_t* lut_dev, uint32_t meta[2],
                uint8_t* inv_host, hipStream_t s);

// Device join entry (hash or sort-merge): pairs of (probe payload, build
// payload); returns the count, writes min(count, cap).
using JoinFn = uint64_t (*)(pimdal_ctx*, const uint32_t*, const uint32_t*, uint64_t, const uint32_t*,
                            const uint32_t*, uint64_t, uint32_t*, uint32_t*, uint64_t, hipStream_t);

// Column forms with validity bitmaps (nulls.hip): SQL null semantics over
// the non-nullable kernels.
uint64_t rows_with_bit(pimdal_ctx* c, const void* bits, uint64_t n, bool want, uint32_t* out_rows, hipStream_t s);
void select_cols_nullable(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* payload, int64_t lo, int64_t hi,
                          pimdal_col** out_payload, pimdal_col** out_rowid);
void join_cols_nullable(JoinFn fn, pimdal_ctx* c, const pimdal_col* bk, const pimdal_col* bp, const pimdal_col* pk,
                        const pimdal_col* pp, pimdal_col** out_p, pimdal_col** out_b);
void aggr_cols_nullable(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* val, pimdal_col** ok, pimdal_col** os,
                        pimdal_col** oc);

// TPC-H Q3 shape (q3.hip): returns the group count, the first min(count,
// cap) groups written in TPC-H order.
uint64_t tpch_q3(pimdal_ctx* c, const pimdal_q3_tables& t, uint32_t segment, int32_t date, const pimdal_q3_result& out,
                 uint64_t cap, hipStream_t s);
uint64_t tpch_q4(pimdal_ctx* c, const pimdal_q4_tables& t, int32_t date_start, int32_t date_end, uint8_t* out_prio,
                 int64_t* out_count, uint64_t cap, hipStream_t s);
uint64_t tpch_q5(pimdal_ctx* c, const pimdal_q5_tables& t, uint32_t region_name, int32_t date_start, int32_t date_end,
                 int32_t* out_nation, int64_t* out_rev, uint64_t cap, hipStream_t s);

}  // namespace pd
