// A reference-style caller: builds arrow::Tables the way
// reference/db_operators builds them and calls the drop-in functions of
// include/pimdal_arrow.hpp instead of the reference's.  Writes each result
// column as raw little-endian bytes into the directory given as argv[1].
#include <cstdio>
#include <cstring>
#include <string>

#include "pimdal_arrow.hpp"
#include "pimdal_gen.h"

template <typename T>
static std::shared_ptr<arrow::Table> make(const std::vector<std::string>& names,
                                          const std::vector<std::vector<T>>& cols,
                                          std::shared_ptr<arrow::DataType> type) {
    arrow::FieldVector f;
    arrow::ArrayVector a;
    for (size_t i = 0; i < cols.size(); i++) {
        auto buf = *arrow::AllocateBuffer(cols[i].size() * sizeof(T));
        memcpy(buf->mutable_data(), cols[i].data(), cols[i].size() * sizeof(T));
        std::shared_ptr<arrow::Buffer> b = std::move(buf);
        a.push_back(arrow::MakeArray(arrow::ArrayData::Make(type, cols[i].size(), {nullptr, b})));
        f.push_back(arrow::field(names[i], type, false));
    }
    return arrow::Table::Make(arrow::schema(f), a);
}

static void dump(const std::string& dir, const std::string& name, const std::shared_ptr<arrow::Table>& t, int col) {
    auto arr = t->column(col)->chunk(0);
    const int w = arr->type()->byte_width();
    FILE* f = fopen((dir + "/" + name).c_str(), "wb");
    fwrite(arr->data()->buffers[1]->data() + arr->offset() * w, w, arr->length(), f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const uint64_t n = 1 << 20;
    std::vector<int32_t> key(n);
    std::vector<uint32_t> ukey(n), inner(n / 4), outer(n);
    const uint64_t s1 = pd_stream(0xD0, 1), s2 = pd_stream(0xD0, 2), s3 = pd_stream(0xD0, 3), s4 = pd_stream(0xD0, 4);
    for (uint64_t i = 0; i < n; i++) {
        key[i] = 1 + (int32_t)pd_uniform_span(pd_rand(s1, i), 50);
        ukey[i] = (uint32_t)(pd_rand(s2, i) >> 32);
        outer[i] = 1 + (uint32_t)pd_uniform_span(pd_rand(s4, i), n / 4);
    }
    for (uint64_t i = 0; i < n / 4; i++) inner[i] = 1 + pd_perm(s3, n / 4, i);
    try {
        auto sel = pimdal_gpu::select(make<int32_t>({"key"}, {key}, arrow::int32()), "key");
        dump(dir, "select_key", sel, 0);
        dump(dir, "select_row", sel, 1);
        auto srt = pimdal_gpu::sort(make<uint32_t>({"key"}, {ukey}, arrow::uint32()), "key");
        dump(dir, "sort_key", srt, 0);
        auto hj = pimdal_gpu::hash_join(make<uint32_t>({"key"}, {inner}, arrow::uint32()),
                                        make<uint32_t>({"key"}, {outer}, arrow::uint32()), "key");
        dump(dir, "hj_inner", hj, 0);
        dump(dir, "hj_outer", hj, 1);
        auto smj = pimdal_gpu::sort_merge_join(make<uint32_t>({"key"}, {inner}, arrow::uint32()),
                                               make<uint32_t>({"key"}, {outer}, arrow::uint32()), "key");
        dump(dir, "smj_inner", smj, 0);
        dump(dir, "smj_outer", smj, 1);
        std::vector<int32_t> val(n);
        for (uint64_t i = 0; i < n; i++) val[i] = 1 + (int32_t)(i % 10);
        auto t = make<int32_t>({"key", "val"}, {key, val}, arrow::int32());
        auto agg = pimdal_gpu::hash_aggr(t, "val");
        dump(dir, "agg_key", agg, 0);
        dump(dir, "agg_sum", agg, 1);
        dump(dir, "agg_cnt", agg, 2);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    printf("dropin ok\n");
    return 0;
}
