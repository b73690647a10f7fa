/*
 * pimdal_parquet.h — C ABI of libpimdal_parquet.so: Parquet columns read
 * straight into device columns of libpimdal_gpu.so.
 *
 * Replaces the reference's table reader, parquet_to_table
 * (pimdal/main/tpc_h/reader/read_table.cpp:18-34: one Parquet file opened
 * with parquet::arrow::OpenFile, the selected columns read as an Arrow
 * table), and the date_to_int packing its callers apply to date literals
 * (read_table.cpp:8-16).  The library is C++ over Arrow/Parquet C++ (the
 * pyarrow wheel's bundled libarrow / libparquet); its entry points take
 * plain pointers and return 0 or a negative PIMDAL_E_* code, with
 * pimdal_parquet_last_error() giving a thread-local message.
 */
#ifndef PIMDAL_PARQUET_H
#define PIMDAL_PARQUET_H

#include <stdint.h>
#include "pimdal_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* date32 columns become the reference's packed dates on the device. */
#define PIMDAL_PARQUET_PACK_DATES 1u
/* String columns become dictionary codes: the distinct values sorted
 * (bytewise), code = rank; uint8 codes for up to 256 values, uint32
 * above.  Without this flag string columns are refused. */
#define PIMDAL_PARQUET_DICT_CODES 2u

const char* pimdal_parquet_last_error(void);

/* Column and row counts of a Parquet file. */
int pimdal_parquet_shape(const char* path, uint32_t* n_columns, uint64_t* n_rows);

/* Name of column `index` (NUL-terminated into buf of buf_len bytes;
 * PIMDAL_E_CAPACITY when it does not fit). */
int pimdal_parquet_column_name(const char* path, uint32_t index, char* buf, uint64_t buf_len);

/* Reads one column (by name, or by index when name is NULL: the sel_col of
 * read_table.cpp:18-34) of every row group into one device column: fixed-
 * width types as pimdal_col_import_chunks takes them (validity kept),
 * date32 packed when PIMDAL_PARQUET_PACK_DATES, strings as codes when
 * PIMDAL_PARQUET_DICT_CODES.  For a string column, dict (may be NULL)
 * receives the values in code order, each NUL-terminated, in dict_bytes
 * bytes (PIMDAL_E_CAPACITY when too small; *n_dict = value count either
 * way).  *out is a new column owned by the caller (pimdal_col_free). */
int pimdal_parquet_read_column(pimdal_ctx* ctx, const char* path, const char* name, int32_t index, uint32_t flags,
                               pimdal_col** out, char* dict, uint64_t dict_bytes, uint32_t* n_dict);

#ifdef __cplusplus
}
#endif

#endif /* PIMDAL_PARQUET_H */
