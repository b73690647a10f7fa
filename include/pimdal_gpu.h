/*
 * pimdal_gpu.h — C ABI of the MI355X (gfx950) DB-operator library.
 *
 * Drop-in for PIMDAL's operator entry points on one GPU.  The reference
 * exposes its operators as C++ functions over arrow::Table
 * (reference/db_operators/select.cpp:77, sort.cpp:42, hash_join.cpp:159,
 * sort_merge_join.cpp:105, aggregate.cpp:95) and as DPU kernels over MRAM
 * argument structs (pimdal/select/sel.h:7-21, pimdal/sort/sort.h:7-25,
 * pimdal/join/hash_join.h:6-82, pimdal/join/sort_merge.h:6-23,
 * pimdal/aggregate/aggregate.h:7-25), with Arrow columns moved to the
 * devices by pimdal/support/transfer_helper.h:87-283.  This header replaces
 * all three with plain pointers and sizes:
 *
 *   - columns move in and out through the Arrow C Data Interface
 *     (pimdal_col_import / pimdal_col_export) or raw host buffers;
 *   - every operator has a column-level form (library allocates the
 *     outputs, like the reference's functions returning a new Table) and a
 *     device-pointer form (`*_dev`: caller passes device buffers, the form
 *     the DPU kernels' argument structs correspond to).
 *
 * Conventions: every function returns 0 (PIMDAL_OK) or a negative
 * PIMDAL_E_* code; pimdal_last_error() gives a thread-local message.  No C++
 * exception crosses this boundary.  Work is issued on the context's stream;
 * functions that return a count or allocate outputs synchronise it.  A
 * context is used by one host thread at a time.
 */
#ifndef PIMDAL_GPU_H
#define PIMDAL_GPU_H

#include <stdint.h>
#include "pimdal_arrow_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIMDAL_OK 0
#define PIMDAL_E_INVALID (-1)      /* bad argument */
#define PIMDAL_E_UNSUPPORTED (-2)  /* e.g. nullable or multi-chunk Arrow input */
#define PIMDAL_E_NOMEM (-3)        /* HBM budget or device allocation exhausted */
#define PIMDAL_E_HIP (-4)          /* HIP runtime error */
#define PIMDAL_E_CAPACITY (-5)     /* output larger than the caller's buffer */
#define PIMDAL_E_INTERNAL (-6)

typedef struct pimdal_ctx pimdal_ctx;
typedef struct pimdal_col pimdal_col;

typedef enum {
    PIMDAL_U8 = 1,
    PIMDAL_I32 = 2,
    PIMDAL_U32 = 3,
    PIMDAL_I64 = 4,
    PIMDAL_U64 = 5,
    PIMDAL_I128 = 6  /* Arrow decimal128(38, 0), little-endian two's complement */
} pimdal_type;

/* ---------------- context ----------------
 * Replaces dpu_alloc / dpu_load and the global transfer contexts
 * (pimdal/support/transfer_helper.h:20-22; pimdal/main/select/host/
 * host_select_sync.cpp:42-45). */
const char* pimdal_last_error(void);
int pimdal_version(void);
int pimdal_ctx_create(int device, uint64_t hbm_budget_bytes, pimdal_ctx** out);
int pimdal_ctx_destroy(pimdal_ctx* ctx);
int pimdal_ctx_sync(pimdal_ctx* ctx);
/* The context's hipStream_t (default priority). */
int pimdal_ctx_stream(pimdal_ctx* ctx, void** stream);
/* Budget accounting: bytes live, reserved (live + cached), peak, budget. */
int pimdal_ctx_mem_stats(pimdal_ctx* ctx, uint64_t* live, uint64_t* reserved, uint64_t* peak, uint64_t* budget);
int pimdal_ctx_set_budget(pimdal_ctx* ctx, uint64_t hbm_budget_bytes);
/* Per-kernel hipEvent timing on the stream each kernel ran on. */
int pimdal_ctx_set_profiling(pimdal_ctx* ctx, int on);
int pimdal_ctx_kernel_stat(pimdal_ctx* ctx, const char* kernel, uint64_t* launches, double* total_ms);
int pimdal_ctx_reset_stats(pimdal_ctx* ctx);
/* Tuning knobs by name (unknown keys are stored and ignored):
 *   sort_block (256, 512, 1024), sort_items (keys per thread: 16 or 8;
 *   1024-thread blocks take 8), sort_early_lookback (0, 1), sort_rank (0 LDS
 *   match masks, 1 per-bit ballots), sort_lbw (look-back statuses read per
 *   round trip: 1, 2, 4, 8, 16, 32), sort_lb_split (1: all threads of a 512-thread
 *   block take part in the look-back, two per digit), sort_nt (non-temporal
 *   pass loads 1, stores 2, both 3), sort_xcd_tiles (S > 0: persistent
 *   chained passes on a co-resident grid, each XCD taking runs of S
 *   consecutive tiles; default 16; 0: one block per tile; if the runtime
 *   refuses the co-resident grid, that pass and the later ones take one
 *   block per tile; sort_xcd_refuse_pass j makes pass j's launch count as
 *   refused, for tests), sort_scheme (0 chained passes with decoupled
 *   look-back, 1 counted passes: per-tile digit counts and a scan, then a
 *   reorder with no look-back), sort_rts_order (tile order of counted
 *   passes: 0 block order, 1 ticket, 2 XCD-contiguous), sort_bucket (1, the
 *   default: for 4 varying digits whose buckets of the two high digits
 *   hold 3072 to 6144 rows by a histogram estimate, as 2^28 uniform keys
 *   do, global passes over the two high digits only, then one pass that
 *   sorts each bucket by the two low digits in LDS; 2: that scheme for any
 *   3 or 4 digits whose estimate fits; 0: one global pass per digit),
 *   sort_bucket_count (1, the default: for 32-bit keys the bucket pass
 *   ranks rows by one LDS atomic into 2048 coarse bins and orders each bin
 *   by (low 16 bits, row), chunks with a bin above 32 rows going to the
 *   match-mask ranking; 0: the match-mask ranking for every chunk);
 *   select_mode (0 one pass with decoupled look-back, 1 count pass + write
 *   pass), select_block (256, 512, 1024 threads; 1024 rows per wave per tile),
 *   select_lookback_width (1 or 4 status words per lane), select_blocks_per_cu;
 *   hj_mode (-1 by build size, 0 partitioned, 1 global table),
 *   hj_global_max_build, hj_probe (1 direct output slots + match bitmap,
 *   0 compacting rounds), hj_block (compacting probe block: 1024 threads x 8
 *   rows or 512 x 16), hj_fanout, hj_rows_per_part, hj_chunk;
 *   hj_onepass (1: up to 2^24 build rows, one partition pass per side into at
 *   most 1024 partitions probed through bucketed LDS tables; 0: 8-bit radix
 *   partition passes), hj_overlap (1: the build side's partition pass runs
 *   on a second stream of the context beside the probe side's; 0: both on
 *   the context stream), hj_csr_rows (8 or 16 probe rows per thread per
 *   round), pw_items (16 or 8 rows per thread in the one-pass partition);
 *   smj_merge (1 direct output slots + match bitmap, 0 ticketed merge with
 *   decoupled look-back over per-tile match counts);
 *   gb_groups_hint, gb_rows_per_part, gb_lds_max_rows, gb_packed,
 *   gb_q1_slots (1024 or 2048), gb_q1_fused (1: Q1 by key filters and
 *   packs rows inside its first hash partition pass; 0: a pack map, then
 *   the partition passes);
 *   aggr_small (1: hash and sort aggregation first try per-block LDS tables
 *   for up to 4096 keys per block and 32768 in all; 0: skip that path),
 *   aggr_direct (1: hash and sort aggregation first try per-block LDS arrays
 *   indexed by key when a sample of the first 2^16 keys spans at most 8192
 *   values and every key falls in the window around it; 0: skip that path),
 *   aggr_dense_max (keys spanning at most this many values, default and
 *   largest 2^24: one range-partition pass into 8192-key partitions summed
 *   in key-indexed LDS arrays, groups in key order; 0: skip that path),
 *   aggr_dense_sample (1: that path's key window from 2^16 keys sampled
 *   across the input, widened, the exact min/max only when a key falls
 *   outside it; 0: the exact min/max of every key);
 *   bloom_groups (1: probes with 64 to 1024 filters hold them in LDS in
 *   groups of 64, each group's blocks reading every probe key; 0: every
 *   filter bit read through L2), bloom_group_min_rows (smallest probe side
 *   that takes the groups, default 2^16), bloom_group_shape (1: 32768-row
 *   chunks, the next one loading; 0: 16384-row chunks; 2: 16384-row
 *   chunks, two loading ahead);
 *   serve_index_load_pct. */
int pimdal_ctx_set_tune(pimdal_ctx* ctx, const char* key, int64_t value);
int pimdal_ctx_get_tune(pimdal_ctx* ctx, const char* key, int64_t* value);
/* Forgets a knob: the built-in default applies again. */
int pimdal_ctx_clear_tune(pimdal_ctx* ctx, const char* key);
/* Autotuner: times candidate knob values for `op` ("sort", "select",
 * "hash_join", "groupby", "q1": the filter + group-by by key, its partition
 * fan-out gb_rows_per_part and table size gb_q1_slots) on synthetic device
 * data of n rows, keeps the
 * fastest in the context, and appends them to cache_path (may be NULL).
 * No counterpart in the reference (its sizes and fan-outs are compile-time
 * CMake definitions, pimdal/main/CMakeLists.txt:11-19). */
int pimdal_autotune(pimdal_ctx* ctx, const char* op, uint64_t n, const char* cache_path);
int pimdal_ctx_load_tuning(pimdal_ctx* ctx, const char* cache_path);

/* ---------------- device memory ---------------- */
int pimdal_dev_alloc(pimdal_ctx* ctx, uint64_t bytes, void** dptr);
int pimdal_dev_free(pimdal_ctx* ctx, void* dptr);
/* Device-to-device copy on the context stream; asynchronous. */
int pimdal_dev_copy(pimdal_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Host <-> device copies through pinned staging; synchronous. */
int pimdal_memcpy_h2d(pimdal_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int pimdal_memcpy_d2h(pimdal_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---------------- columns ----------------
 * Replaces dist_table / get_buf (pimdal/support/transfer_helper.h:207-236):
 * a column is one device-resident, fixed-width array, with an optional
 * validity bitmap in Arrow's layout (bit i % 8 of byte i / 8, 1 = valid).
 * The column forms of the operators (pimdal_select, pimdal_hash_join,
 * pimdal_sort_merge_join, pimdal_hash_aggr, pimdal_sort_aggr) give NULLs
 * SQL semantics, which the reference's non-nullable operators lack: a NULL
 * key satisfies no select predicate and joins nothing; payload validity
 * follows the kept rows and the joined pairs; in an aggregation the
 * NULL-key rows form one group (the last, key NULL), SUM skips NULL values
 * (NULL for a group with none) and COUNT counts the non-NULL values
 * (groups then come in ascending key order).  The `*_dev` forms take no
 * validity. */
int pimdal_col_alloc(pimdal_ctx* ctx, pimdal_type type, uint64_t length, pimdal_col** out);
int pimdal_col_from_host(pimdal_ctx* ctx, pimdal_type type, const void* data, uint64_t length, pimdal_col** out);
/* Wraps caller-owned device memory (not freed by pimdal_col_free). */
int pimdal_col_wrap(pimdal_ctx* ctx, pimdal_type type, void* dptr, uint64_t length, pimdal_col** out);
int pimdal_col_to_host(pimdal_ctx* ctx, const pimdal_col* col, void* dst, uint64_t dst_bytes);
int pimdal_col_info(const pimdal_col* col, pimdal_type* type, uint64_t* length, void** dptr);
int pimdal_col_free(pimdal_ctx* ctx, pimdal_col* col);
/* The column's device validity bitmap (NULL when it has no NULLs) and its
 * NULL count. */
int pimdal_col_validity(const pimdal_col* col, void** dptr, uint64_t* null_count);
/* Sets the validity from a host bitmap of ceil(length / 8) bytes (Arrow
 * layout); NULL, or a bitmap with every bit set, removes it. */
int pimdal_col_set_validity(pimdal_ctx* ctx, pimdal_col* col, const uint8_t* bits);
/* Arrow C Data Interface.  Import borrows the array until it returns,
 * copies its validity bitmap (offset applied) when it has NULLs, and
 * rejects non-primitive or unsupported types with PIMDAL_E_UNSUPPORTED.
 * Export allocates host buffers owned by `out` (a validity buffer when the
 * column has NULLs); the caller calls out->release and
 * out_schema->release. */
int pimdal_col_import(pimdal_ctx* ctx, const struct ArrowArray* array, const struct ArrowSchema* schema, pimdal_col** out);
/* Chunked column (an arrow::ChunkedArray, as a Parquet reader returns one
 * chunk per row group): the `nchunks` arrays, all of `schema`'s type and
 * none with nulls, concatenated in order into one device column.  Widens
 * the single-chunk input of the reference operators
 * (reference/db_operators/select.cpp:22 takes chunk(0)) and of dist_table
 * (pimdal/support/transfer_helper.h:87-283).  Every chunk is checked
 * before anything is copied.  The Arrow date32 format ("tdD", int32 days
 * since 1970-01-01) imports as PIMDAL_I32. */
int pimdal_col_import_chunks(pimdal_ctx* ctx, const struct ArrowArray* const* chunks, uint32_t nchunks,
                             const struct ArrowSchema* schema, pimdal_col** out);
int pimdal_col_export(pimdal_ctx* ctx, const pimdal_col* col, struct ArrowArray* out, struct ArrowSchema* out_schema);

/* ---------------- dates ----------------
 * The packed date of pimdal/main/tpc_h/reader/read_table.cpp:8-16
 * (date_to_int): day + (month << 5) + ((year - 1900) << 9), the form the
 * TPC-H programs compare their date columns and literals in
 * (pimdal/main/tpc_h/query6/host/host_q6.cpp:46-47, query1/host/host_q1.cpp:49).
 * pimdal_date_pack: from numbers (month 1..12, day 1..31, as strptime's
 * %m and %d accept them; the day is not checked against the month).
 * pimdal_date_pack_str: from "YYYY-MM-DD" (a 1-4 digit year, 1-2 digit month
 * and day; surrounding blanks allowed), the literal form date_to_int takes;
 * a malformed string is PIMDAL_E_INVALID (the reference leaves the result
 * undefined).
 * pimdal_date32_pack_dev: a device column of Arrow date32 values (int32 days
 * since 1970-01-01, proleptic Gregorian) to packed dates, n values,
 * in place allowed (out == days).  Asynchronous. */
int pimdal_date_pack(int32_t year, int32_t month, int32_t day, int32_t* out);
int pimdal_date_pack_str(const char* yyyy_mm_dd, int32_t* out);
int pimdal_date32_pack_dev(pimdal_ctx* ctx, const int32_t* days, uint64_t n, int32_t* out);

/* ---------------- selection ----------------
 * reference/db_operators/select.cpp:77 select(table, column) and
 * pimdal/select/sel.h:21 sel_kernel.  Keeps rows with lo <= key < hi (the
 * predicate form of select.cpp:33-36, bounds as arguments), order
 * preserving.  Outputs the payload and the row index of every kept row;
 * pass the key column as payload to get the reference's compacted keys.
 * key type: PIMDAL_U8, PIMDAL_U32, PIMDAL_I32 or PIMDAL_I64 (compared as
 * signed 64-bit against lo/hi); the payload is 4 bytes wide.  Without a
 * payload column only row ids are output.
 * out_payload / out_rowid may be NULL to skip them. */
int pimdal_select(pimdal_ctx* ctx, const pimdal_col* key, const pimdal_col* payload, int64_t lo, int64_t hi,
                  pimdal_col** out_payload, pimdal_col** out_rowid);
/* Device form: out buffers hold n rows; *d_count (device) receives the
 * count.  Asynchronous. */
int pimdal_select_dev(pimdal_ctx* ctx, pimdal_type key_type, const void* key, const uint32_t* payload, uint64_t n,
                      int64_t lo, int64_t hi, uint32_t* out_payload, uint32_t* out_rowid, uint64_t* d_count);
/* No-allocation form of pimdal_select_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, outputs and result as pimdal_select_dev
 * (always the one-pass kernel); its look-back status words live in a
 * static device array of the library (1 MiB + 256 B per GPU, held until
 * the process ends), so it calls no hipMalloc, hipFree or allocating
 * routine.  Limit: n <= 2^29 rows, else PIMDAL_E_CAPACITY.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (two streams,
 * two contexts, two host threads) race and must be serialised by the
 * caller.  Asynchronous. */
int pimdal_select_dev_noalloc(pimdal_ctx* ctx, pimdal_type key_type, const void* key, const uint32_t* payload,
                              uint64_t n, int64_t lo, int64_t hi, uint32_t* out_payload, uint32_t* out_rowid,
                              uint64_t* d_count);

/* ---------------- sort ----------------
 * reference/db_operators/sort.cpp:42 sort(table, column) and
 * pimdal/sort/sort.h:20-25 sort_kernel.  Stable LSD radix sort of
 * (uint32 key, payload) pairs; payload width 4 (PIMDAL_U32/I32) or 8
 * (PIMDAL_U64/I64) bytes, or NULL for keys only.  Bits that are equal in
 * every key are not sorted on (detected on the device). */
int pimdal_sort_pairs(pimdal_ctx* ctx, const pimdal_col* key, const pimdal_col* payload,
                      pimdal_col** out_key, pimdal_col** out_payload);
/* Device form.  Sorts bits [begin_bit, end_bit) of the key; pass 0, 32 for
 * the whole key, or begin_bit = end_bit = -1 to detect the varying bits
 * (one device->host read).  val_width 0, 4 or 8.  Asynchronous unless
 * detecting bits. */
int pimdal_sort_pairs_dev(pimdal_ctx* ctx, const uint32_t* keys_in, const void* vals_in, uint32_t val_width,
                          uint64_t n, uint32_t* keys_out, void* vals_out, int begin_bit, int end_bit);

/* ---------------- joins ----------------
 * Hash join: reference/db_operators/hash_join.cpp:159 hash_join(inner,
 * outer, column) and pimdal/join/hash_join.h:62-82 (hash_kernel,
 * part_kernel, merge_kernel).  Sort-merge join: reference/db_operators/
 * sort_merge_join.cpp:105 sort_merge_join(inner, outer, column) and
 * pimdal/join/sort_merge.h:23 merge_kernel.  Equi-join on uint32 keys
 * (int32 keys reinterpreted: equality is the same).  Build = inner,
 * probe = outer.  Every matching pair is emitted as (probe payload, build
 * payload); duplicate build keys are allowed.  A NULL payload means the
 * key itself.  Hash-join order is unspecified; sort-merge order is
 * (key, probe row, build row).
 * Device form: out buffers hold out_capacity rows; *out_count (host)
 * receives the number of pairs even when it exceeds out_capacity (then
 * PIMDAL_E_CAPACITY is returned and the first out_capacity pairs are
 * written).  Synchronous. */
int pimdal_hash_join(pimdal_ctx* ctx, const pimdal_col* build_key, const pimdal_col* build_payload,
                     const pimdal_col* probe_key, const pimdal_col* probe_payload,
                     pimdal_col** out_probe_payload, pimdal_col** out_build_payload);
int pimdal_hash_join_dev(pimdal_ctx* ctx, const uint32_t* build_key, const uint32_t* build_payload, uint64_t n_build,
                         const uint32_t* probe_key, const uint32_t* probe_payload, uint64_t n_probe,
                         uint32_t* out_probe_payload, uint32_t* out_build_payload, uint64_t out_capacity,
                         uint64_t* out_count);
int pimdal_sort_merge_join(pimdal_ctx* ctx, const pimdal_col* build_key, const pimdal_col* build_payload,
                           const pimdal_col* probe_key, const pimdal_col* probe_payload,
                           pimdal_col** out_probe_payload, pimdal_col** out_build_payload);
int pimdal_sort_merge_join_dev(pimdal_ctx* ctx, const uint32_t* build_key, const uint32_t* build_payload,
                               uint64_t n_build, const uint32_t* probe_key, const uint32_t* probe_payload,
                               uint64_t n_probe, uint32_t* out_probe_payload, uint32_t* out_build_payload,
                               uint64_t out_capacity, uint64_t* out_count);

/* ---------------- Bloom-filter semi-join ----------------
 * The PIM library's filter_kernel / match_kernel (pimdal/join/hash_join.c:
 * 117-152, 237-341; insert_bloom_filter / probe_bloom_filter,
 * pimdal/hash/hash_steps.c:143-183; mixers pimdal/hash/hash_func.h:5-58).
 * Keys fall into 2^part_bits partitions by bits [shift - part_bits, shift)
 * of hash0(key) (part_kernel's MSB-first radix partition, whose `shift` is
 * this argument; part_bits <= 16, part_bits <= shift <= 32).  Each
 * partition has a 16384-bit filter (2048 bytes, the reference's byte
 * layout) in which hash1..hash4(key) % 16384 are set.
 * pimdal_bloom_build_dev: the filters of n build keys into `filters`
 * (2^part_bits x 2048 bytes, overwritten).
 * pimdal_bloom_probe_dev: the probe rows whose four bits are all set in
 * their partition's filter: every row whose key is a build key, plus the
 * filter's false positives, exactly as the reference's filters decide.
 * Writes their payload (the key when payload is NULL) and row index in row
 * order (the reference emits them partition by partition); *d_count
 * (device) receives the count.  out_payload / out_rowid hold n rows or are
 * NULL.  pimdal_bloom_semi_join_dev: both, with the filters in scratch
 * memory.  Asynchronous. */
int pimdal_bloom_build_dev(pimdal_ctx* ctx, const uint32_t* keys, uint64_t n, uint32_t part_bits, uint32_t shift,
                           uint32_t* filters);
int pimdal_bloom_probe_dev(pimdal_ctx* ctx, const uint32_t* keys, const uint32_t* payload, uint64_t n,
                           uint32_t part_bits, uint32_t shift, const uint32_t* filters, uint32_t* out_payload,
                           uint32_t* out_rowid, uint64_t* d_count);
int pimdal_bloom_semi_join_dev(pimdal_ctx* ctx, const uint32_t* build_keys, uint64_t nb, const uint32_t* probe_keys,
                               const uint32_t* probe_payload, uint64_t np, uint32_t part_bits, uint32_t shift,
                               uint32_t* out_payload, uint32_t* out_rowid, uint64_t* d_count);

/* ---------------- aggregation ----------------
 * Hash aggregation: reference/db_operators/aggregate.cpp:95 hash_aggr(table,
 * column) and pimdal/aggregate/aggregate.h:25 group_kernel (hash variant,
 * pimdal/aggregate/aggregate_hash.c:123-193).  Sort aggregation: the sort
 * variant of group_kernel (pimdal/aggregate/aggregate.c:82-158).  Group by a
 * uint32 key; per group SUM(val) (exact int64; the reference sums into
 * int32) and COUNT.  Hash output order is unspecified; sort output is in
 * ascending key order.  Device form: outputs hold `capacity` groups;
 * *n_groups (host) receives the count.  Synchronous. */
int pimdal_hash_aggr_dev(pimdal_ctx* ctx, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                         int64_t* out_sum, int64_t* out_count, uint64_t capacity, uint64_t* n_groups);
int pimdal_sort_aggr_dev(pimdal_ctx* ctx, const uint32_t* key, const int64_t* val, uint64_t n, uint32_t* out_key,
                         int64_t* out_sum, int64_t* out_count, uint64_t capacity, uint64_t* n_groups);
int pimdal_hash_aggr(pimdal_ctx* ctx, const pimdal_col* key, const pimdal_col* val, pimdal_col** out_key,
                     pimdal_col** out_sum, pimdal_col** out_count);
int pimdal_sort_aggr(pimdal_ctx* ctx, const pimdal_col* key, const pimdal_col* val, pimdal_col** out_key,
                     pimdal_col** out_sum, pimdal_col** out_count);

/* Fused filter + group-by with the TPC-H Q1 measures:
 * pimdal/main/tpc_h/query1/dpu/kernel_q1_1.c:59-228 (filter on the date,
 * gather, disc_price = price*(100-disc), charge = disc_price*(100+tax)),
 * kernel_q1_2.c:37-46 (per-group sums, count) and host_q1.cpp:233-330.
 * Keeps rows with date <= cutoff.  Groups by (flag1, flag2) when key is
 * NULL (group key = flag1*256 + flag2; output in ascending key order), or
 * by the int32 key column (unspecified order).  Sums are exact: int64, and
 * sum_charge as int128 (two int64 words per group, low word first). */
typedef struct {
    int32_t* key;
    int64_t* sum_qty;
    int64_t* sum_price;
    int64_t* sum_disc;
    int64_t* sum_disc_price;
    int64_t* sum_charge; /* 2 words per group */
    int64_t* count;
} pimdal_q1_result;
int pimdal_filter_groupby_q1_dev(pimdal_ctx* ctx, const int32_t* date, const uint8_t* flag1, const uint8_t* flag2,
                                 const int32_t* key, const int64_t* quantity, const int64_t* price,
                                 const int64_t* discount, const int64_t* tax, uint64_t n, int32_t cutoff,
                                 const pimdal_q1_result* out, uint64_t capacity, uint64_t* n_groups);

/* TPC-H Q6 shape: pimdal/main/tpc_h/query6/dpu/kernel_q6.c:91-141 (the
 * predicates pred_date, pred_quantity and pred_discount applied in turn,
 * then price * discount and an int64 sum) with the parameters of
 * host_q6.cpp:46-49 (one year of dates, discount 6 +- 1, quantity < 24)
 * and the sum over devices of host_q6.cpp:60-66.  Keeps rows with
 * date_lo <= date < date_hi, quantity < qty_lt and disc_lo <= discount <=
 * disc_hi; revenue = sum of price * discount over them, int64 modulo 2^64
 * like the reference's int64 accumulators (the reference prints it divided
 * by 10^4).  n_rows = rows kept.  Synchronises the stream. */
int pimdal_q6_dev(pimdal_ctx* ctx, const int32_t* date, const int64_t* quantity, const int64_t* price,
                  const int64_t* discount, uint64_t n, int32_t date_lo, int32_t date_hi, int64_t disc_lo,
                  int64_t disc_hi, int64_t qty_lt, int64_t* revenue, uint64_t* n_rows);

/* Asynchronous forms of the two calls above, for plans and for callers that
 * keep results on the device: nothing is read back to the host, so they
 * can be captured into a plan (pimdal_plan_begin).  Results go to device
 * words: *d_n_groups (uint64) and *d_status (int32): 0, or
 * PIMDAL_E_CAPACITY when there are more groups than `capacity` (the count
 * is still exact, `capacity` groups are written), or PIMDAL_E_UNSUPPORTED
 * when the data leaves what the asynchronous form covers: more than 8
 * (flag1, flag2) combinations (flag form), or a measure beyond the packed
 * widths (price beyond int32, quantity beyond int16, discount or tax beyond
 * int8) or a hash partition whose groups overflow its LDS table (key form).
 * Then the outputs are undefined and pimdal_filter_groupby_q1_dev gives the
 * result.  The flag form's groups come out in ascending key order, as in
 * pimdal_filter_groupby_q1_dev; the key form's in no particular order.
 * pimdal_q6_dev_async leaves the revenue (int64) and kept-row count
 * (uint64) in device words. */
int pimdal_filter_groupby_q1_dev_async(pimdal_ctx* ctx, const int32_t* date, const uint8_t* flag1,
                                       const uint8_t* flag2, const int32_t* key, const int64_t* quantity,
                                       const int64_t* price, const int64_t* discount, const int64_t* tax, uint64_t n,
                                       int32_t cutoff, const pimdal_q1_result* out, uint64_t capacity,
                                       uint64_t* d_n_groups, int32_t* d_status);
int pimdal_q6_dev_async(pimdal_ctx* ctx, const int32_t* date, const int64_t* quantity, const int64_t* price,
                        const int64_t* discount, uint64_t n, int32_t date_lo, int32_t date_hi, int64_t disc_lo,
                        int64_t disc_hi, int64_t qty_lt, int64_t* d_revenue, uint64_t* d_rows);

/* TPC-H Q3 shape: pimdal/main/tpc_h/query3/host/host_q3.cpp:471-519 with
 * its kernels kernel_q3_1.c:116-137 (customers of a market segment),
 * kernel_q3_2.c:146-190 (orders with o_orderdate < date), kernel_q3_3.c
 * (join on custkey), kernel_q3_4.c:120-141 (lineitems with l_shipdate >
 * date), kernel_q3_5.c (join on orderkey, revenue = l_extendedprice *
 * (100 - l_discount) summed per (orderkey, orderdate, shippriority), top
 * rows by revenue).  Dates are the packed dates of pimdal_date_pack; the
 * segment is a dictionary code (the index of e.g. "BUILDING" in the
 * c_mktsegment dictionary; the reference matches 16-byte strings).
 * o_orderkey must be unique (PIMDAL_E_INVALID otherwise).  Writes the
 * first min(groups, capacity) groups in TPC-H order (revenue descending,
 * o_orderdate, then orderkey ascending); *n_groups = all groups.
 * Revenue sums wrap modulo 2^64 like the reference's int64.  Synchronous. */
typedef struct {
    const int32_t* c_custkey;
    const uint8_t* c_mktsegment;
    uint64_t n_customer;
    const int32_t* o_orderkey;
    const int32_t* o_custkey;
    const int32_t* o_orderdate;
    const int32_t* o_shippriority;
    uint64_t n_orders;
    const int32_t* l_orderkey;
    const int64_t* l_extendedprice;
    const int64_t* l_discount;
    const int32_t* l_shipdate;
    uint64_t n_lineitem;
} pimdal_q3_tables;
typedef struct {
    int32_t* orderkey;
    int64_t* revenue;
    int32_t* orderdate;
    int32_t* shippriority;
} pimdal_q3_result;
int pimdal_tpch_q3_dev(pimdal_ctx* ctx, const pimdal_q3_tables* tables, uint32_t segment, int32_t date,
                       const pimdal_q3_result* out, uint64_t capacity, uint64_t* n_groups);

/* TPC-H Q4 shape: pimdal/main/tpc_h/query4/host/host_q4.cpp:303-330 with its
 * kernels kernel_q4_1.c:161-201 (orders with date_start <= o_orderdate <
 * date_end), kernel_q4_2.c:104-140 (lineitems with l_commitdate <
 * l_receiptdate), kernel_q4_3.c:144-209 (join on orderkey, one row per
 * order) and kernel_q4_4.c:72-99 with host_q4.cpp:211-231 (orders counted
 * per o_orderpriority, ordered by priority).  Dates are packed dates; the
 * priority is a dictionary code (the reference groups 16-byte strings; a
 * sorted dictionary keeps their order).  o_orderkey must be unique among
 * the qualifying orders (PIMDAL_E_INVALID otherwise).  Writes the first
 * min(groups, capacity) (priority, order count) rows in priority order;
 * *n_groups = priorities with at least one order.  Synchronous. */
typedef struct {
    const int32_t* o_orderkey;
    const int32_t* o_orderdate;
    const uint8_t* o_orderpriority;
    uint64_t n_orders;
    const int32_t* l_orderkey;
    const int32_t* l_commitdate;
    const int32_t* l_receiptdate;
    uint64_t n_lineitem;
} pimdal_q4_tables;
int pimdal_tpch_q4_dev(pimdal_ctx* ctx, const pimdal_q4_tables* tables, int32_t date_start, int32_t date_end,
                       uint8_t* out_priority, int64_t* out_count, uint64_t capacity, uint64_t* n_groups);

/* TPC-H Q5 shape: pimdal/main/tpc_h/query5/host/host_q5.cpp:566-640 with its
 * kernels kernel_q5_1.c (the region of r_name, its nations, their
 * customers), kernel_q5_2.c (orders with date_start <= o_orderdate <
 * date_end), kernel_q5_3.c (customers join orders on custkey),
 * kernel_q5_4.c / kernel_q5_5.c (orders join lineitems on orderkey),
 * kernel_q5_6.c / kernel_q5_7.c (join suppliers on suppkey and nation,
 * revenue = l_extendedprice * (100 - l_discount) summed per nation) and
 * host_q5.cpp:459-500 (sums over the devices, revenue descending).  The
 * region name is a dictionary code (the reference matches the substring
 * "ASIA" in 16-byte names); nation keys must lie in [0, 1024)
 * (PIMDAL_E_UNSUPPORTED otherwise); custkey, orderkey and suppkey must be
 * unique among the qualifying customers, orders and the suppliers
 * (PIMDAL_E_INVALID).  Writes the first min(groups, capacity) (nationkey,
 * revenue) rows, revenue descending then nationkey ascending; *n_groups =
 * nations with at least one matching lineitem.  Revenue sums wrap modulo
 * 2^64 like the reference's int64.  Synchronous. */
typedef struct {
    const int32_t* r_regionkey;
    const uint8_t* r_name;
    uint64_t n_region;
    const int32_t* n_nationkey;
    const int32_t* n_regionkey;
    uint64_t n_nation;
    const int32_t* c_custkey;
    const int32_t* c_nationkey;
    uint64_t n_customer;
    const int32_t* o_orderkey;
    const int32_t* o_custkey;
    const int32_t* o_orderdate;
    uint64_t n_orders;
    const int32_t* l_orderkey;
    const int32_t* l_suppkey;
    const int64_t* l_extendedprice;
    const int64_t* l_discount;
    uint64_t n_lineitem;
    const int32_t* s_suppkey;
    const int32_t* s_nationkey;
    uint64_t n_supplier;
} pimdal_q5_tables;
int pimdal_tpch_q5_dev(pimdal_ctx* ctx, const pimdal_q5_tables* tables, uint32_t region_name, int32_t date_start,
                       int32_t date_end, int32_t* out_nationkey, int64_t* out_revenue, uint64_t capacity,
                       uint64_t* n_groups);

/* ---------------- captured plans ----------------
 * A plan records a sequence of asynchronous operator calls on the context's
 * stream as one hipGraph and replays it: the query plans of
 * pimdal/main/tpc_h/query1/host/host_q1.cpp:345-380 (filter + aggregation
 * kernels launched one after another, the host waiting on each) become one
 * graph launch.  Between pimdal_plan_begin and pimdal_plan_end the calls
 * run as captured work, not at once.  Capturable: pimdal_select_dev,
 * pimdal_select_dev_noalloc, pimdal_sort_pairs_dev with an explicit bit
 * range, pimdal_date32_pack_dev, pimdal_bloom_build_dev,
 * pimdal_bloom_probe_dev, pimdal_bloom_semi_join_dev, pimdal_dev_copy,
 * pimdal_filter_groupby_q1_dev_async and pimdal_q6_dev_async.  Calls that
 * wait for the device or return host results (joins, aggregations that
 * return counts, column forms, host copies, pimdal_ctx_sync) are refused
 * with PIMDAL_E_UNSUPPORTED inside a capture, and pimdal_plan_end then
 * fails too (nothing to launch; destroy the plan).  While capturing, the
 * radix passes use ordinary launches instead of cooperative ones and the
 * scratch memory of every captured call is kept by the plan (charged to
 * the context's budget) until pimdal_plan_destroy.  Launches are
 * asynchronous on the context stream; they read the inputs and write the
 * outputs at the addresses given at capture time.  Destroy plans before
 * their context. */
typedef struct pimdal_plan pimdal_plan;
int pimdal_plan_begin(pimdal_ctx* ctx, pimdal_plan** out);
int pimdal_plan_end(pimdal_plan* plan);
int pimdal_plan_launch(pimdal_plan* plan);
int pimdal_plan_destroy(pimdal_plan* plan);

/* ---------------- serving ----------------
 * No counterpart in the reference (its drivers run one operator per
 * binary, pimdal/main/<operator>/host): a request loop over tables kept resident in
 * HBM.  Queries of three kinds run over a slice [offset, offset + rows) of
 * the fact table:
 *   PIMDAL_Q_SELECT  rows with price < param                -> fact row ids
 *   PIMDAL_Q_JOIN    slice orderkey = dim key, dim date < param -> (fact row, dim row)
 *   PIMDAL_Q_GROUPBY date <= param, group by (flag1, flag2), Q1 measures
 * Each result is reported as a row count and an order-independent checksum
 * (include/pimdal_gen.h: pd_ck_select / pd_ck_join / pd_ck_group).
 * PIMDAL_SERVE_SEQUENTIAL runs one query at a time through the plain
 * operators; PIMDAL_SERVE_BATCHED coalesces arrived queries into batches on
 * `slots` streams (continuous batching, one hipGraph replay per batch).
 * The select checksum depends on the output order (each row id with its
 * output position); join and group-by checksums do not. */
typedef struct pimdal_server pimdal_server;

enum { PIMDAL_Q_SELECT = 1, PIMDAL_Q_JOIN = 2, PIMDAL_Q_GROUPBY = 3 };
enum { PIMDAL_SERVE_SEQUENTIAL = 0, PIMDAL_SERVE_BATCHED = 1 };

typedef struct {
    const int32_t* orderkey;
    const int64_t* quantity;
    const int64_t* price;
    const int64_t* discount;
    const int64_t* tax;
    const int32_t* date;
    const uint8_t* flag1;
    const uint8_t* flag2;
    uint64_t n_fact;
    const int32_t* dim_key;  /* unique */
    const int32_t* dim_date;
    const uint8_t* dim_category;
    uint64_t n_dim;
} pimdal_tables;

typedef struct {
    int32_t kind;
    uint32_t tenant;    /* admission shares the serving heap among the tenants with work */
    uint64_t offset;
    uint64_t rows;
    int64_t param;
    double arrival_us;  /* open-loop arrival, microseconds after the run starts; non-decreasing */
} pimdal_query;

typedef struct {
    uint64_t count;
    uint64_t checksum;
    double start_us, end_us, latency_us;  /* latency = end - arrival */
} pimdal_query_result;

typedef struct {
    double makespan_us, qps, p50_us, p99_us, mean_us;
    uint64_t batches, max_inflight;
    uint64_t heap_bytes, heap_peak_bytes;   /* serving heap and its largest use during the run */
    uint64_t tenants;
} pimdal_serve_stats;

/* Builds the resident dimension-key index (and, unless the context's
 * serve_join_index knob is 0, the foreign-key join index: the dimension row
 * and date of every fact row, read by the joins of both modes), `slots`
 * stream slots, and the serving heap of heap_bytes (0: half of what the
 * context's budget leaves, at most 8 GiB), charged to the budget.  A
 * batched query occupies its footprint in the heap while its batch runs:
 * 4 bytes per slice row (select) or 8 (join) for the outputs, 8 per
 * 16384-row tile for look-back statuses, 512 per 2^16-row chunk (group-by).
 * Batches are admitted while their footprint fits the heap's largest free
 * range; a query larger than the whole heap is refused (PIMDAL_E_NOMEM).
 * Queries of each tenant are admitted in arrival order, tenants in turn;
 * a tenant holds at most heap / (tenants with work) while others wait. */
int pimdal_server_create(pimdal_ctx* ctx, const pimdal_tables* tables, int slots, uint64_t heap_bytes,
                         pimdal_server** out);
int pimdal_server_run(pimdal_server* server, const pimdal_query* queries, uint64_t n, int mode,
                      pimdal_query_result* results, pimdal_serve_stats* stats);
int pimdal_server_destroy(pimdal_server* server);

/* ---------------- synthetic data (device generators) ----------------
 * Counter-based generators, identical to include/pimdal_gen.h on the host;
 * they replace reference/db_operators/shared.cpp:9-60. */
int pimdal_gen_u32(pimdal_ctx* ctx, uint64_t seed, uint32_t column, uint64_t n, uint32_t* d_out);
int pimdal_gen_range(pimdal_ctx* ctx, pimdal_type type, uint64_t seed, uint32_t column, uint64_t n,
                     int64_t lo, int64_t hi, void* d_out);
int pimdal_gen_perm(pimdal_ctx* ctx, uint64_t seed, uint32_t column, uint64_t m, uint64_t offset, uint64_t n,
                    uint32_t* d_out);
/* d_out[i] = perm_m(d_in[i] + delta) */
int pimdal_perm_apply(pimdal_ctx* ctx, uint64_t seed, uint32_t column, uint64_t m, uint64_t n, int64_t delta,
                      const uint32_t* d_in, uint32_t* d_out);
int pimdal_gen_iota(pimdal_ctx* ctx, uint64_t n, uint32_t start, uint32_t* d_out);
int pimdal_gen_zipf(pimdal_ctx* ctx, uint64_t seed, uint32_t column, uint64_t m, double s, uint64_t n,
                    uint32_t* d_out_rank);
int pimdal_fmix32(pimdal_ctx* ctx, uint64_t n, const uint32_t* d_in, uint32_t* d_out);
int pimdal_gen_orderkey(pimdal_ctx* ctx, uint64_t seed, uint64_t n, uint64_t groups, int32_t* d_out);
/* Dimension keys: the `groups` distinct fact orderkey values, each once, in a
 * seeded order. */
int pimdal_gen_dim_keys(pimdal_ctx* ctx, uint64_t seed, uint64_t groups, int32_t* d_out);

#ifdef __cplusplus
}
#endif

#endif /* PIMDAL_GPU_H */
