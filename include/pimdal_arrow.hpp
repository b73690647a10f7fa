// pimdal_arrow.hpp — header-only Arrow C++ adapter over the C ABI.
//
// The same function shapes as the reference's CPU operators, so a caller of
// reference/db_operators can switch by changing the namespace:
//   select(table, column)                 reference/db_operators/select.cpp:77
//   sort(table, column)                   reference/db_operators/sort.cpp:42
//   hash_join(inner, outer, column)       reference/db_operators/hash_join.cpp:159
//   sort_merge_join(inner, outer, column) reference/db_operators/sort_merge_join.cpp:105
//   hash_aggr(table, column)              reference/db_operators/aggregate.cpp:95
// Columns cross through Arrow's C Data Interface (arrow/c/bridge.h) into
// device-resident columns and back.  Differences from the reference, all
// documented in the C header: the select predicate bounds are arguments
// (the reference hard-codes 10 <= x < 30), the join returns every matching
// pair including duplicate build keys, and hash_aggr sums into int64 and adds
// a count column.  Errors throw std::runtime_error with the library message.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <arrow/api.h>
#include <arrow/c/bridge.h>

#include "pimdal_gpu.h"

namespace pimdal_gpu {

inline void check(int rc) {
    if (rc != PIMDAL_OK) throw std::runtime_error(std::string("pimdal: ") + pimdal_last_error());
}

// One device context for the calling thread (device 0, budget: 90 % of free HBM).
inline pimdal_ctx* context() {
    struct Holder {
        pimdal_ctx* c = nullptr;
        Holder() { check(pimdal_ctx_create(0, 0, &c)); }
        ~Holder() { pimdal_ctx_destroy(c); }
    };
    static thread_local Holder h;
    return h.c;
}

struct Col {
    pimdal_col* p = nullptr;
    ~Col() { if (p) pimdal_col_free(context(), p); }
};

inline std::shared_ptr<arrow::Array> single_chunk(const std::shared_ptr<arrow::Table>& t, const std::string& name) {
    auto cc = t->GetColumnByName(name);
    if (!cc) throw std::runtime_error("pimdal: no column " + name);
    if (cc->num_chunks() != 1) throw std::runtime_error("pimdal: column " + name + " must have one chunk");
    return cc->chunk(0);
}

inline void to_device(const std::shared_ptr<arrow::Array>& a, Col& out) {
    struct ArrowArray ca;
    struct ArrowSchema cs;
    auto st = arrow::ExportArray(*a, &ca, &cs);
    if (!st.ok()) throw std::runtime_error(st.ToString());
    int rc = pimdal_col_import(context(), &ca, &cs, &out.p);
    ca.release(&ca);
    cs.release(&cs);
    check(rc);
}

inline std::shared_ptr<arrow::Array> to_arrow(const Col& c) {
    struct ArrowArray ca;
    struct ArrowSchema cs;
    check(pimdal_col_export(context(), c.p, &ca, &cs));
    auto r = arrow::ImportArray(&ca, &cs);
    if (!r.ok()) throw std::runtime_error(r.status().ToString());
    return *r;
}

inline std::shared_ptr<arrow::Table> table(const std::vector<std::string>& names, const std::vector<const Col*>& cols) {
    arrow::FieldVector f;
    arrow::ArrayVector a;
    for (size_t i = 0; i < cols.size(); i++) {
        a.push_back(to_arrow(*cols[i]));
        f.push_back(arrow::field(names[i], a.back()->type(), false));
    }
    return arrow::Table::Make(arrow::schema(f), a);
}

// Rows of `table` with lo <= column < hi: Table{key, row}.
inline std::shared_ptr<arrow::Table> select(std::shared_ptr<arrow::Table> t, std::string column, int64_t lo,
                                            int64_t hi) {
    Col key, out, rows;
    to_device(single_chunk(t, column), key);
    check(pimdal_select(context(), key.p, key.p, lo, hi, &out.p, &rows.p));
    return table({"key", "row"}, {&out, &rows});
}

// The reference predicate (select.cpp:33-36): 10 <= x < 30.
inline std::shared_ptr<arrow::Table> select(std::shared_ptr<arrow::Table> t, std::string column) {
    return select(std::move(t), std::move(column), 10, 30);
}

inline std::shared_ptr<arrow::Table> sort(std::shared_ptr<arrow::Table> t, std::string column) {
    Col key, out, unused;
    to_device(single_chunk(t, column), key);
    check(pimdal_sort_pairs(context(), key.p, nullptr, &out.p, nullptr));
    return table({"key"}, {&out});
}

inline std::shared_ptr<arrow::Table> join_impl(bool hash, std::shared_ptr<arrow::Table> inner,
                                               std::shared_ptr<arrow::Table> outer, const std::string& column) {
    Col ik, ok, oi, oo;
    to_device(single_chunk(inner, column), ik);
    to_device(single_chunk(outer, column), ok);
    if (hash)
        check(pimdal_hash_join(context(), ik.p, nullptr, ok.p, nullptr, &oo.p, &oi.p));
    else
        check(pimdal_sort_merge_join(context(), ik.p, nullptr, ok.p, nullptr, &oo.p, &oi.p));
    return table({"inner_key", "outer_key"}, {&oi, &oo});
}

inline std::shared_ptr<arrow::Table> hash_join(std::shared_ptr<arrow::Table> inner, std::shared_ptr<arrow::Table> outer,
                                               std::string column) {
    return join_impl(true, std::move(inner), std::move(outer), column);
}

inline std::shared_ptr<arrow::Table> sort_merge_join(std::shared_ptr<arrow::Table> inner,
                                                     std::shared_ptr<arrow::Table> outer, std::string column) {
    return join_impl(false, std::move(inner), std::move(outer), column);
}

// Group by "key", SUM(column) as int64 plus COUNT: Table{key, val, count}.
inline std::shared_ptr<arrow::Table> hash_aggr(std::shared_ptr<arrow::Table> t, std::string column) {
    Col key, val, ok, os, oc;
    to_device(single_chunk(t, "key"), key);
    auto v = single_chunk(t, column);
    if (v->type_id() != arrow::Type::INT64) {
        auto r = v->View(arrow::int64());  // only for 8-byte types; widen 32-bit values below
        std::shared_ptr<arrow::Array> w;
        if (v->type_id() == arrow::Type::INT32) {
            arrow::Int64Builder b;
            auto i32 = std::static_pointer_cast<arrow::Int32Array>(v);
            if (!b.Reserve(i32->length()).ok()) throw std::runtime_error("pimdal: allocation");
            for (int64_t i = 0; i < i32->length(); i++) b.UnsafeAppend(i32->Value(i));
            if (!b.Finish(&w).ok()) throw std::runtime_error("pimdal: allocation");
        } else if (r.ok()) {
            w = *r;
        } else {
            throw std::runtime_error("pimdal: aggregate column must be int32 or int64");
        }
        v = w;
    }
    to_device(v, val);
    check(pimdal_hash_aggr(context(), key.p, val.p, &ok.p, &os.p, &oc.p));
    return table({"key", "val", "count"}, {&ok, &os, &oc});
}

}  // namespace pimdal_gpu
