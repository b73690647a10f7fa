// parquet_reader.cpp — libpimdal_parquet.so: Parquet columns into device
// columns.
//
// Reference: pimdal/main/tpc_h/reader/read_table.cpp:18-34
// (parquet_to_table: ReadableFile::Open, parquet::arrow::OpenFile,
// ReadTable of the selected columns) and :8-16 (date_to_int).  Here one
// column is read at a time (ReadColumn: one Arrow chunk per row group), its
// chunks are handed through the Arrow C Data Interface to
// pimdal_col_import_chunks (pinned, double-buffered staging into one device
// column, validity kept), date32 columns are packed on the device
// (pimdal_date32_pack_dev) and string columns become sorted dictionary
// codes on the host.  Host C++ over Arrow / Parquet C++; no device code.
#include <arrow/api.h>
#include <arrow/c/bridge.h>
#include <arrow/io/api.h>
#include <parquet/arrow/reader.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pimdal_parquet.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

// An open file reader, or the Arrow error as a message.
int open_reader(const char* path, std::unique_ptr<parquet::arrow::FileReader>* out) {
    if (!path) return fail(PIMDAL_E_INVALID, "path is NULL");
    auto in = arrow::io::ReadableFile::Open(path);
    if (!in.ok()) return fail(PIMDAL_E_INVALID, "cannot open " + std::string(path) + ": " + in.status().ToString());
    auto r = parquet::arrow::OpenFile(*in, arrow::default_memory_pool());
    if (!r.ok()) return fail(PIMDAL_E_INVALID, "not a Parquet file " + std::string(path) + ": " + r.status().ToString());
    *out = std::move(*r);
    return PIMDAL_OK;
}

// A library error: its code, with the library's message.
int lib_fail(int rc, const char* what) {
    return fail(rc, std::string(what) + ": " + (pimdal_last_error() ? pimdal_last_error() : ""));
}

// String values as sorted dictionary codes (uint8 up to 256 values, else
// uint32) with the validity of the source, into a new device column.
template <typename StringArrayT>
int dict_codes(pimdal_ctx* ctx, const arrow::ChunkedArray& col, pimdal_col** out, char* dict, uint64_t dict_bytes,
               uint32_t* n_dict) {
    std::map<std::string, uint32_t> values;
    const uint64_t n = (uint64_t)col.length();
    for (const auto& ch : col.chunks()) {
        const auto& a = static_cast<const StringArrayT&>(*ch);
        for (int64_t i = 0; i < a.length(); i++)
            if (a.IsValid(i)) values.emplace(std::string(a.GetView(i)), 0u);
    }
    uint32_t code = 0;
    uint64_t need = 0;
    for (auto& kv : values) {
        kv.second = code++;
        need += kv.first.size() + 1;
    }
    if (n_dict) *n_dict = (uint32_t)values.size();
    if (dict) {
        if (need > dict_bytes) return fail(PIMDAL_E_CAPACITY, "dictionary buffer too small");
        char* p = dict;
        for (const auto& kv : values) {
            memcpy(p, kv.first.data(), kv.first.size());
            p[kv.first.size()] = '\0';
            p += kv.first.size() + 1;
        }
    }
    const bool narrow = values.size() <= 256;
    std::vector<uint8_t> c8(narrow ? n : 0);
    std::vector<uint32_t> c32(narrow ? 0 : n);
    std::vector<uint8_t> bits((n + 7) / 8, 0);
    uint64_t r = 0, nulls = 0;
    for (const auto& ch : col.chunks()) {
        const auto& a = static_cast<const StringArrayT&>(*ch);
        for (int64_t i = 0; i < a.length(); i++, r++) {
            uint32_t v = 0;
            if (a.IsValid(i)) {
                v = values.at(std::string(a.GetView(i)));
                bits[r >> 3] |= (uint8_t)(1u << (r & 7));
            } else {
                nulls++;
            }
            if (narrow) c8[r] = (uint8_t)v;
            else c32[r] = v;
        }
    }
    pimdal_col* c = nullptr;
    int rc = narrow ? pimdal_col_from_host(ctx, PIMDAL_U8, c8.data(), n, &c)
                    : pimdal_col_from_host(ctx, PIMDAL_U32, c32.data(), n, &c);
    if (rc) return lib_fail(rc, "dictionary codes to the device");
    if (nulls && (rc = pimdal_col_set_validity(ctx, c, bits.data()))) {
        pimdal_col_free(ctx, c);
        return lib_fail(rc, "validity of the dictionary codes");
    }
    *out = c;
    return PIMDAL_OK;
}

}  // namespace

extern "C" const char* pimdal_parquet_last_error(void) { return g_error.c_str(); }

extern "C" int pimdal_parquet_shape(const char* path, uint32_t* n_columns, uint64_t* n_rows) {
    try {
        std::unique_ptr<parquet::arrow::FileReader> rd;
        if (int rc = open_reader(path, &rd)) return rc;
        std::shared_ptr<arrow::Schema> schema;
        auto st = rd->GetSchema(&schema);
        if (!st.ok()) return fail(PIMDAL_E_INVALID, st.ToString());
        if (n_columns) *n_columns = (uint32_t)schema->num_fields();
        if (n_rows) *n_rows = (uint64_t)rd->parquet_reader()->metadata()->num_rows();
        return PIMDAL_OK;
    } catch (const std::exception& e) {
        return fail(PIMDAL_E_INTERNAL, e.what());
    }
}

extern "C" int pimdal_parquet_column_name(const char* path, uint32_t index, char* buf, uint64_t buf_len) {
    try {
        if (!buf) return fail(PIMDAL_E_INVALID, "buf is NULL");
        std::unique_ptr<parquet::arrow::FileReader> rd;
        if (int rc = open_reader(path, &rd)) return rc;
        std::shared_ptr<arrow::Schema> schema;
        auto st = rd->GetSchema(&schema);
        if (!st.ok()) return fail(PIMDAL_E_INVALID, st.ToString());
        if ((int)index >= schema->num_fields()) return fail(PIMDAL_E_INVALID, "column index out of range");
        const std::string& name = schema->field((int)index)->name();
        if (name.size() + 1 > buf_len) return fail(PIMDAL_E_CAPACITY, "name buffer too small");
        memcpy(buf, name.c_str(), name.size() + 1);
        return PIMDAL_OK;
    } catch (const std::exception& e) {
        return fail(PIMDAL_E_INTERNAL, e.what());
    }
}

extern "C" int pimdal_parquet_read_column(pimdal_ctx* ctx, const char* path, const char* name, int32_t index,
                                          uint32_t flags, pimdal_col** out, char* dict, uint64_t dict_bytes,
                                          uint32_t* n_dict) {
    try {
        if (!ctx || !out) return fail(PIMDAL_E_INVALID, "NULL argument");
        std::unique_ptr<parquet::arrow::FileReader> rd;
        if (int rc = open_reader(path, &rd)) return rc;
        std::shared_ptr<arrow::Schema> schema;
        auto st = rd->GetSchema(&schema);
        if (!st.ok()) return fail(PIMDAL_E_INVALID, st.ToString());
        int i = index;
        if (name) {
            i = schema->GetFieldIndex(name);
            if (i < 0) return fail(PIMDAL_E_INVALID, "no column named " + std::string(name));
        }
        if (i < 0 || i >= schema->num_fields()) return fail(PIMDAL_E_INVALID, "column index out of range");
        std::shared_ptr<arrow::ChunkedArray> col;
        st = rd->ReadColumn(i, &col);
        if (!st.ok()) return fail(PIMDAL_E_INVALID, "reading column: " + st.ToString());
        const auto type = col->type();
        if (type->id() == arrow::Type::STRING || type->id() == arrow::Type::LARGE_STRING) {
            if (!(flags & PIMDAL_PARQUET_DICT_CODES))
                return fail(PIMDAL_E_UNSUPPORTED, "string column: pass PIMDAL_PARQUET_DICT_CODES for its codes");
            return type->id() == arrow::Type::STRING
                       ? dict_codes<arrow::StringArray>(ctx, *col, out, dict, dict_bytes, n_dict)
                       : dict_codes<arrow::LargeStringArray>(ctx, *col, out, dict, dict_bytes, n_dict);
        }
        // fixed-width chunks through the C Data Interface, one device column
        const int k = col->num_chunks();
        std::vector<struct ArrowArray> arrays((size_t)std::max(k, 1));
        std::vector<const struct ArrowArray*> ptrs((size_t)std::max(k, 1));
        struct ArrowSchema sch;
        st = arrow::ExportType(*type, &sch);
        if (!st.ok()) return fail(PIMDAL_E_UNSUPPORTED, "column type: " + st.ToString());
        int exported = 0, rc = PIMDAL_OK;
        for (int c = 0; c < k && rc == PIMDAL_OK; c++) {
            st = arrow::ExportArray(*col->chunk(c), &arrays[(size_t)c]);
            if (!st.ok()) rc = fail(PIMDAL_E_UNSUPPORTED, "column chunk: " + st.ToString());
            else ptrs[(size_t)c] = &arrays[(size_t)c], exported++;
        }
        pimdal_col* res = nullptr;
        if (rc == PIMDAL_OK) {
            rc = pimdal_col_import_chunks(ctx, ptrs.data(), (uint32_t)k, &sch, &res);
            if (rc) rc = lib_fail(rc, "importing the column");
        }
        for (int c = 0; c < exported; c++)
            if (arrays[(size_t)c].release) arrays[(size_t)c].release(&arrays[(size_t)c]);
        if (sch.release) sch.release(&sch);
        if (rc) return rc;
        if ((flags & PIMDAL_PARQUET_PACK_DATES) && type->id() == arrow::Type::DATE32) {
            uint64_t n = 0;
            void* d = nullptr;
            pimdal_col_info(res, nullptr, &n, &d);
            if ((rc = pimdal_date32_pack_dev(ctx, (const int32_t*)d, n, (int32_t*)d)) ||
                (rc = pimdal_ctx_sync(ctx))) {
                pimdal_col_free(ctx, res);
                return lib_fail(rc, "packing the dates");
            }
        }
        if (n_dict) *n_dict = 0;
        *out = res;
        return PIMDAL_OK;
    } catch (const std::exception& e) {
        return fail(PIMDAL_E_INTERNAL, e.what());
    }
}
