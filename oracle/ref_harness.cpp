// ref_harness.cpp — C entry points over the reference's own CPU operators.
//
// TEST INFRASTRUCTURE ONLY.  Each build of this file includes one
// reference/db_operators/*.cpp source, as it lies under the reference tree
// (the path comes from the build recipe oracle/ref.mk), with the reference's
// main() renamed away, and wraps its operator function for ctypes.  The
// results pin the CPU restatement in oracle.cpp.  The reference compiles
// BUFFER_SIZE in as a constant (NUM_BLOCKS = BUFFER_SIZE / 8000 blocks are
// processed), so each wrapper requires n == REF_N.
#include <stdint.h>
#include <string.h>

#define main pimdal_ref_unused_main
#include REF_SRC
#undef main

namespace {

template <typename T>
std::shared_ptr<arrow::Table> make_table(const T* data, uint64_t n, std::shared_ptr<arrow::DataType> type) {
    auto buf = *arrow::AllocateBuffer(n * sizeof(T));
    memcpy(buf->mutable_data(), data, n * sizeof(T));
    std::shared_ptr<arrow::Buffer> b = std::move(buf);
    auto arr = arrow::MakeArray(arrow::ArrayData::Make(type, n, {nullptr, b}));
    return arrow::Table::Make(arrow::schema({arrow::field("key", type, false)}), {arr});
}

template <typename T>
uint64_t copy_col(const std::shared_ptr<arrow::Table>& t, int col, T* out, uint64_t cap) {
    auto a = t->column(col)->chunk(0);
    const T* v = a->data()->GetValues<T>(1);
    uint64_t n = std::min<uint64_t>(cap, a->length());
    memcpy(out, v, n * sizeof(T));
    return a->length();
}

}  // namespace

#ifdef REF_SELECT
// reference/db_operators/select.cpp:77 select(table, "key"), predicate 10 <= x < 30
extern "C" uint64_t ref_select_i32(const int32_t* key, uint64_t n, int32_t* out, uint64_t cap) {
    if (n != REF_N) return ~0ull;
    auto res = select(make_table(key, n, arrow::int32()), "key");
    return copy_col(res, 0, out, cap);
}
#endif

#ifdef REF_SORT
// reference/db_operators/sort.cpp:42 sort(table, "key")
extern "C" uint64_t ref_sort_u32(const uint32_t* key, uint64_t n, uint32_t* out) {
    if (n != REF_N) return ~0ull;
    auto res = sort(make_table(key, n, arrow::uint32()), "key");
    return copy_col(res, 0, out, n);
}
#endif

#ifdef REF_HASH_JOIN
// reference/db_operators/hash_join.cpp:159 hash_join(inner, outer, "key")
extern "C" uint64_t ref_hash_join_u32(const uint32_t* inner, const uint32_t* outer, uint64_t n, uint32_t* out_inner,
                                      uint32_t* out_outer) {
    if (n != REF_N) return ~0ull;
    auto res = hash_join(make_table(inner, n, arrow::uint32()), make_table(outer, n, arrow::uint32()), "key");
    copy_col(res, 0, out_inner, n);
    return copy_col(res, 1, out_outer, n);
}
#endif

#ifdef REF_AGGREGATE
// reference/db_operators/aggregate.cpp:95 hash_aggr(table, "val")
extern "C" uint64_t ref_hash_aggr(const uint32_t* key, const int32_t* val, uint64_t n, uint32_t* out_key,
                                  int32_t* out_val, uint64_t cap) {
    if (n != REF_N) return ~0ull;
    auto kb = *arrow::AllocateBuffer(n * 4);
    memcpy(kb->mutable_data(), key, n * 4);
    auto vb = *arrow::AllocateBuffer(n * 4);
    memcpy(vb->mutable_data(), val, n * 4);
    std::shared_ptr<arrow::Buffer> kbs = std::move(kb), vbs = std::move(vb);
    auto ka = arrow::MakeArray(arrow::ArrayData::Make(arrow::uint32(), n, {nullptr, kbs}));
    auto va = arrow::MakeArray(arrow::ArrayData::Make(arrow::int32(), n, {nullptr, vbs}));
    auto t = arrow::Table::Make(arrow::schema({arrow::field("key", arrow::uint32(), false),
                                               arrow::field("val", arrow::int32(), false)}), {ka, va});
    auto res = hash_aggr(t, "val");
    copy_col(res, 0, out_key, cap);
    return copy_col(res, 1, out_val, cap);
}
#endif
