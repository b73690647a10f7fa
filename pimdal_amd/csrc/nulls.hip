// nulls.hip — SQL null semantics for the column-level operators.
//
// The reference operators take non-nullable columns only (they read
// chunk(0)->GetValues<T>(1) and ignore validity, reference/db_operators/
// select.cpp:22, hash_join.cpp:165-170, aggregate.cpp:100-103).  Columns
// here may carry an Arrow validity bitmap, and the column forms give them
// the SQL meaning:
//  * select: a NULL key satisfies no predicate; a kept row's payload keeps
//    its validity;
//  * joins: a NULL key equals nothing (neither side's NULL-key rows join);
//    payload validity follows each output pair;
//  * aggregation: the rows with a NULL key form one group (reported last,
//    its key NULL); SUM skips NULL values (NULL when a group has none);
//    COUNT counts the non-NULL values.
// The device paths stay the non-nullable kernels: the rows to use are
// compacted by their validity bits first (order kept), the operator runs on
// them with row ids as payloads, and values and validity are gathered
// through those row ids afterwards.
#include <algorithm>
#include <vector>

#include "pd_runtime.hpp"
#include "pd_device.hpp"
#include "pd_ops.hpp"

namespace pd {

constexpr uint32_t NB_CHUNK = 2048;   // rows per wave: 64 validity words

__device__ __forceinline__ uint32_t nb_word(const uint32_t* __restrict__ bits, uint64_t n, uint64_t wd, bool want) {
    if (wd * 32 >= n) return 0u;
    uint32_t v = want ? bits[wd] : ~bits[wd];
    const uint64_t left = n - wd * 32;
    if (left < 32) v &= (1u << left) - 1u;   // rows past n
    return v;
}

// Rows per 2048-row chunk whose bit equals `want`.
__global__ __launch_bounds__(256) void k_nb_count(const uint32_t* __restrict__ bits, uint64_t n, bool want,
                                                  uint32_t* __restrict__ counts, uint64_t nchunks) {
    const uint64_t ch = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / PD_WAVE;
    if (ch >= nchunks) return;
    const uint32_t c = pd::wave_sum(__popc(nb_word(bits, n, ch * 64 + pd::lane(), want)));
    if (pd::lane() == 0) counts[ch] = c;
}

// The rows themselves, in row order, from the chunks' offsets.
__global__ __launch_bounds__(256) void k_nb_rows(const uint32_t* __restrict__ bits, uint64_t n, bool want,
                                                 const uint32_t* __restrict__ offs, uint64_t nchunks,
                                                 uint32_t* __restrict__ rows) {
    const uint64_t ch = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / PD_WAVE;
    if (ch >= nchunks) return;
    const uint64_t wd = ch * 64 + pd::lane();
    uint32_t v = nb_word(bits, n, wd, want);
    const uint32_t c = __popc(v);
    uint64_t o = offs[ch] + pd::wave_incl_scan(c) - c;
    while (v) {
        const int j = __builtin_ctz(v);
        v &= v - 1;
        rows[o++] = (uint32_t)(wd * 32 + (uint64_t)j);
    }
}

template <typename T>
__global__ void k_nb_gather(const T* __restrict__ src, const uint32_t* __restrict__ rows, uint64_t k,
                            T* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = src[rows[i]];
}

// Validity of gathered rows: output bit i = source bit rows[i]; one ballot
// per 64 outputs (two words); valid rows counted with one atomic per block.
__global__ __launch_bounds__(256) void k_nb_gather_bits(const uint32_t* __restrict__ src, const uint32_t* __restrict__ rows,
                                                        uint64_t k, uint32_t* __restrict__ out,
                                                        unsigned long long* __restrict__ valid) {
    __shared__ uint32_t s_w[4];
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    bool b = false;
    if (i < k) {
        const uint32_t r = rows[i];
        b = (src[r >> 5] >> (r & 31)) & 1;
    }
    const uint64_t bal = pd::ballot(b);
    const uint64_t wv = i / PD_WAVE;
    if (pd::lane() == 0 && wv * PD_WAVE < k) {
        out[2 * wv] = (uint32_t)bal;
        if ((2 * wv + 1) * 32 < k) out[2 * wv + 1] = (uint32_t)(bal >> 32);
    }
    if (pd::lane() == 0) s_w[threadIdx.x / PD_WAVE] = pd::popc64(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t c = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (c) atomicAdd(valid, (unsigned long long)c);
    }
}

// SUM / COUNT inputs: val0 = value or 0 when NULL, nn = 1 when not NULL.
__global__ void k_nb_split(const int64_t* __restrict__ val, const uint32_t* __restrict__ bits, uint64_t n,
                           int64_t* __restrict__ val0, int64_t* __restrict__ nn) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool ok = !bits || ((bits[i >> 5] >> (i & 31)) & 1);
        val0[i] = ok ? val[i] : 0;
        nn[i] = ok ? 1 : 0;
    }
}

// Sums of val0 and nn over `rows` (wrapping), one atomic pair per block.
__global__ __launch_bounds__(256) void k_nb_sum_rows(const int64_t* __restrict__ val0, const int64_t* __restrict__ nn,
                                                     const uint32_t* __restrict__ rows, uint64_t k,
                                                     unsigned long long* __restrict__ out) {
    __shared__ uint64_t s_w[4][2];
    uint64_t a = 0, b = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        a += (uint64_t)val0[rows[i]];
        b += (uint64_t)nn[rows[i]];
    }
    a = pd::wave_sum64(a);
    b = pd::wave_sum64(b);
    if (pd::lane() == 0) {
        s_w[threadIdx.x / PD_WAVE][0] = a;
        s_w[threadIdx.x / PD_WAVE][1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t x = 0, y = 0;
        for (int w = 0; w < 4; w++) { x += s_w[w][0]; y += s_w[w][1]; }
        atomicAdd(&out[0], (unsigned long long)x);
        atomicAdd(&out[1], (unsigned long long)y);
    }
}

// Validity of a count column's groups: bit i = cnt[i] > 0.
__global__ __launch_bounds__(256) void k_nb_nonzero_bits(const int64_t* __restrict__ cnt, uint64_t g,
                                                         uint32_t* __restrict__ out,
                                                         unsigned long long* __restrict__ valid) {
    __shared__ uint32_t s_w[4];
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const bool b = i < g && cnt[i] > 0;
    const uint64_t bal = pd::ballot(b);
    const uint64_t wv = i / PD_WAVE;
    if (pd::lane() == 0 && wv * PD_WAVE < g) {
        out[2 * wv] = (uint32_t)bal;
        if ((2 * wv + 1) * 32 < g) out[2 * wv + 1] = (uint32_t)(bal >> 32);
    }
    if (pd::lane() == 0) s_w[threadIdx.x / PD_WAVE] = pd::popc64(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t c = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (c) atomicAdd(valid, (unsigned long long)c);
    }
}

__global__ void k_nb_iota(uint32_t* __restrict__ out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)i;
}

static dim3 nb_grid(pimdal_ctx* c, uint64_t n) {
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), (uint64_t)c->cu_count * 16)));
}

// Order-preserving list of the rows < n whose validity bit equals `want`.
uint64_t rows_with_bit(pimdal_ctx* c, const void* bits, uint64_t n, bool want, uint32_t* out_rows, hipStream_t s) {
    if (n == 0) return 0;
    const uint64_t nch = ceil_div(n, NB_CHUNK);
    Scratch cnt(c, (nch + 1) * 4, s), tot(c, 8, s);
    const dim3 g((uint32_t)ceil_div(nch * PD_WAVE, 256));
    launch(c, "null_rows", k_nb_count, g, dim3(256), 0, s, (const uint32_t*)bits, n, want, cnt.as<uint32_t>(), nch);
    exclusive_scan_u32(c, cnt.as<uint32_t>(), cnt.as<uint32_t>(), nch, tot.as<uint64_t>(), s);
    launch(c, "null_rows", k_nb_rows, g, dim3(256), 0, s, (const uint32_t*)bits, n, want,
           (const uint32_t*)cnt.as<uint32_t>(), nch, out_rows);
    return read_u64(c, tot.get(), s);
}

template <typename T>
static void gather(pimdal_ctx* c, const void* src, const uint32_t* rows, uint64_t k, void* out, hipStream_t s) {
    launch(c, "null_gather", k_nb_gather<T>, nb_grid(c, k), dim3(256), 0, s, (const T*)src, rows, k, (T*)out);
}

// New pool column of n rows of `type` (length n).
static pimdal_col* new_col(pimdal_ctx* c, pimdal_type t, uint64_t n) {
    auto* col = new pimdal_col{t, n, n, nullptr, true};
    try {
        col->data = c->pool.alloc(std::max<uint64_t>(n, 1) * type_width(t), c->stream);
    } catch (...) {
        delete col;
        throw;
    }
    return col;
}

static void free_col(pimdal_ctx* c, pimdal_col*& col) {
    if (!col) return;
    if (col->owned && col->data) c->pool.free(col->data, c->stream);
    if (col->validity) c->pool.free(col->validity, c->stream);
    delete col;
    col = nullptr;
}

// Gives `out` (k rows) the validity of `src` rows `rows` (when src has any
// NULL); no bitmap when every gathered row is valid.
static void gather_validity(pimdal_ctx* c, const pimdal_col* src, const uint32_t* rows, uint64_t k, pimdal_col* out,
                            hipStream_t s) {
    if (!src || !src->validity || k == 0) return;
    const uint64_t nw = valid_words(k);
    void* bits = c->pool.alloc(std::max<uint64_t>(nw, 1) * 4, s);
    try {
        Scratch valid(c, 8, s);
        fill_u64(c, valid.get(), 0, 1, s);
        launch(c, "null_gather", k_nb_gather_bits, dim3((uint32_t)ceil_div(k, 256)), dim3(256), 0, s,
               (const uint32_t*)src->validity, rows, k, (uint32_t*)bits, valid.as<unsigned long long>());
        const uint64_t v = read_u64(c, valid.get(), s);
        if (v == k) {
            c->pool.free(bits, s);
            return;
        }
        out->validity = bits;
        out->null_count = k - v;
    } catch (...) {
        c->pool.free(bits, s);
        throw;
    }
}

// ---- select ----
void select_cols_nullable(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* payload, int64_t lo, int64_t hi,
                          pimdal_col** out_payload, pimdal_col** out_rowid) {
    hipStream_t s = c->stream;
    const uint64_t n = key->length;
    PD_REQUIRE(payload || key->type != PIMDAL_I64 || !out_payload, PIMDAL_E_INVALID,
               "int64 keys need a 32-bit payload column");
    Scratch rid(c, std::max<uint64_t>(n, 1) * 4, s), cnt(c, 8, s);
    select_range(c, key->type, key->data, nullptr, n, lo, hi, nullptr, rid.as<uint32_t>(), cnt.as<uint64_t>(), s);
    uint64_t k = read_u64(c, cnt.get(), s);
    const uint32_t* rows = rid.as<uint32_t>();
    Scratch rid2(c, key->validity ? std::max<uint64_t>(k, 1) * 4 : 0, s), kb(c, key->validity ? valid_words(k) * 4 + 4 : 0, s),
        pos(c, key->validity ? std::max<uint64_t>(k, 1) * 4 : 0, s);
    if (key->validity && k) {
        // keep the selected rows whose key is not NULL
        Scratch valid(c, 8, s);
        fill_u64(c, valid.get(), 0, 1, s);
        launch(c, "null_gather", k_nb_gather_bits, dim3((uint32_t)ceil_div(k, 256)), dim3(256), 0, s,
               (const uint32_t*)key->validity, rows, k, kb.as<uint32_t>(), valid.as<unsigned long long>());
        const uint64_t k2 = rows_with_bit(c, kb.get(), k, true, pos.as<uint32_t>(), s);
        gather<uint32_t>(c, rows, pos.as<uint32_t>(), k2, rid2.get(), s);
        rows = rid2.as<uint32_t>();
        k = k2;
    }
    pimdal_col* op = nullptr;
    pimdal_col* orow = nullptr;
    try {
        if (out_rowid) {
            orow = new_col(c, PIMDAL_U32, k);
            copy_d2d(c, orow->data, rows, k * 4, s);
        }
        if (out_payload) {
            const pimdal_col* src = payload ? payload : key;
            op = new_col(c, src->type, k);
            gather<uint32_t>(c, src->data, rows, k, op->data, s);
            gather_validity(c, payload, rows, k, op, s);
        }
        PD_HIP(hipStreamSynchronize(s));
    } catch (...) {
        free_col(c, op);
        free_col(c, orow);
        throw;
    }
    if (out_payload) *out_payload = op;
    if (out_rowid) *out_rowid = orow;
}

// ---- joins ----
void join_cols_nullable(JoinFn fn, pimdal_ctx* c, const pimdal_col* bk, const pimdal_col* bp, const pimdal_col* pk,
                        const pimdal_col* pp, pimdal_col** out_p, pimdal_col** out_b) {
    hipStream_t s = c->stream;
    // each side: the rows with a non-NULL key, their keys, their row ids
    auto side = [&](const pimdal_col* k, Scratch& rows, Scratch& keys) -> uint64_t {
        const uint64_t n = k->length;
        if (k->validity) {
            const uint64_t m = rows_with_bit(c, k->validity, n, true, rows.as<uint32_t>(), s);
            gather<uint32_t>(c, k->data, rows.as<uint32_t>(), m, keys.get(), s);
            return m;
        }
        launch(c, "null_rows", k_nb_iota, nb_grid(c, n), dim3(256), 0, s, rows.as<uint32_t>(), n);
        copy_d2d(c, keys.get(), k->data, n * 4, s);
        return n;
    };
    Scratch brow(c, std::max<uint64_t>(bk->length, 1) * 4, s), bkey(c, std::max<uint64_t>(bk->length, 1) * 4, s);
    Scratch prow(c, std::max<uint64_t>(pk->length, 1) * 4, s), pkey(c, std::max<uint64_t>(pk->length, 1) * 4, s);
    const uint64_t nb = side(bk, brow, bkey), np = side(pk, prow, pkey);
    uint64_t cap = std::max<uint64_t>(np, 1);
    Scratch* op = new Scratch(c, cap * 4, s);
    Scratch* ob = new Scratch(c, cap * 4, s);
    pimdal_col* a = nullptr;
    pimdal_col* b = nullptr;
    try {
        uint64_t k = fn(c, bkey.as<uint32_t>(), brow.as<uint32_t>(), nb, pkey.as<uint32_t>(), prow.as<uint32_t>(), np,
                        op->as<uint32_t>(), ob->as<uint32_t>(), cap, s);
        if (k > cap) {
            delete op;
            delete ob;
            op = ob = nullptr;
            cap = k;
            op = new Scratch(c, cap * 4, s);
            ob = new Scratch(c, cap * 4, s);
            k = fn(c, bkey.as<uint32_t>(), brow.as<uint32_t>(), nb, pkey.as<uint32_t>(), prow.as<uint32_t>(), np,
                   op->as<uint32_t>(), ob->as<uint32_t>(), cap, s);
        }
        const pimdal_col* ps = pp ? pp : pk;
        const pimdal_col* bs = bp ? bp : bk;
        a = new_col(c, ps->type, k);
        b = new_col(c, bs->type, k);
        gather<uint32_t>(c, ps->data, op->as<uint32_t>(), k, a->data, s);
        gather<uint32_t>(c, bs->data, ob->as<uint32_t>(), k, b->data, s);
        gather_validity(c, pp, op->as<uint32_t>(), k, a, s);
        gather_validity(c, bp, ob->as<uint32_t>(), k, b, s);
        PD_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete op;
        delete ob;
        free_col(c, a);
        free_col(c, b);
        throw;
    }
    delete op;
    delete ob;
    *out_p = a;
    *out_b = b;
}

// ---- aggregation ----
void aggr_cols_nullable(pimdal_ctx* c, const pimdal_col* key, const pimdal_col* val, pimdal_col** ok, pimdal_col** os,
                        pimdal_col** oc) {
    hipStream_t s = c->stream;
    const uint64_t n = key->length;
    Scratch val0(c, std::max<uint64_t>(n, 1) * 8, s), nn(c, std::max<uint64_t>(n, 1) * 8, s);
    launch(c, "null_split", k_nb_split, nb_grid(c, n), dim3(256), 0, s, (const int64_t*)val->data,
           (const uint32_t*)val->validity, n, val0.as<int64_t>(), nn.as<int64_t>());
    // rows with a non-NULL key
    Scratch rows(c, key->validity ? std::max<uint64_t>(n, 1) * 4 : 0, s);
    Scratch kv(c, key->validity ? std::max<uint64_t>(n, 1) * 4 : 0, s), v0(c, key->validity ? std::max<uint64_t>(n, 1) * 8 : 0, s),
        v1(c, key->validity ? std::max<uint64_t>(n, 1) * 8 : 0, s);
    const uint32_t* keys = (const uint32_t*)key->data;
    const int64_t* a0 = val0.as<int64_t>();
    const int64_t* a1 = nn.as<int64_t>();
    uint64_t nk = n;
    if (key->validity) {
        nk = rows_with_bit(c, key->validity, n, true, rows.as<uint32_t>(), s);
        gather<uint32_t>(c, key->data, rows.as<uint32_t>(), nk, kv.get(), s);
        gather<int64_t>(c, val0.get(), rows.as<uint32_t>(), nk, v0.get(), s);
        gather<int64_t>(c, nn.get(), rows.as<uint32_t>(), nk, v1.get(), s);
        keys = kv.as<uint32_t>();
        a0 = v0.as<int64_t>();
        a1 = v1.as<int64_t>();
    }
    // groups of the non-NULL keys in ascending key order: SUM of the values
    // with NULLs as 0, then (same groups, same order) SUM of the non-NULL flags
    Scratch gk(c, std::max<uint64_t>(nk, 1) * 4, s), gs(c, std::max<uint64_t>(nk, 1) * 8, s),
        gr(c, std::max<uint64_t>(nk, 1) * 8, s), gk2(c, std::max<uint64_t>(nk, 1) * 4, s),
        gn(c, std::max<uint64_t>(nk, 1) * 8, s);
    const uint64_t g = sort_aggr(c, keys, a0, nk, gk.as<uint32_t>(), gs.as<int64_t>(), gr.as<int64_t>(), nk, s);
    sort_aggr(c, keys, a1, nk, gk2.as<uint32_t>(), gn.as<int64_t>(), gr.as<int64_t>(), nk, s);
    const bool null_group = key->validity && nk < n;
    const uint64_t G = g + (null_group ? 1 : 0);
    pimdal_col *ck = nullptr, *cs = nullptr, *cc = nullptr;
    try {
        ck = new_col(c, key->type, G);
        cs = new_col(c, PIMDAL_I64, G);
        cc = new_col(c, PIMDAL_I64, G);
        copy_d2d(c, ck->data, gk.get(), g * 4, s);
        copy_d2d(c, cs->data, gs.get(), g * 8, s);
        copy_d2d(c, cc->data, gn.get(), g * 8, s);
        if (null_group) {
            // the NULL-key rows' group, last: key 0 marked NULL
            Scratch nrows(c, (n - nk) * 4, s), sums(c, 16, s);
            const uint64_t m = rows_with_bit(c, key->validity, n, false, nrows.as<uint32_t>(), s);
            fill_u64(c, sums.get(), 0, 2, s);
            launch(c, "null_sum", k_nb_sum_rows, nb_grid(c, m), dim3(256), 0, s, (const int64_t*)val0.as<int64_t>(),
                   (const int64_t*)nn.as<int64_t>(), (const uint32_t*)nrows.as<uint32_t>(), m,
                   sums.as<unsigned long long>());
            fill_u32(c, (uint32_t*)ck->data + g, 0, 1, s);
            copy_d2d(c, (int64_t*)cs->data + g, sums.get(), 8, s);
            copy_d2d(c, (int64_t*)cc->data + g, sums.as<int64_t>() + 1, 8, s);
            std::vector<uint32_t> kb(valid_words(G), 0xFFFFFFFFu);
            if (G % 32) kb.back() = (1u << (G % 32)) - 1u;
            kb[g >> 5] &= ~(1u << (g & 31));
            set_validity_words(c, ck, kb.data(), 1);
        }
        // SUM is NULL for a group without a non-NULL value
        if (val->validity && G) {
            Scratch valid(c, 8, s);
            fill_u64(c, valid.get(), 0, 1, s);
            void* bits = c->pool.alloc(valid_words(G) * 4, s);
            cs->validity = bits;
            launch(c, "null_bits", k_nb_nonzero_bits, dim3((uint32_t)ceil_div(G, 256)), dim3(256), 0, s,
                   (const int64_t*)cc->data, G, (uint32_t*)bits, valid.as<unsigned long long>());
            const uint64_t v = read_u64(c, valid.get(), s);
            if (v == G) {
                c->pool.free(bits, s);
                cs->validity = nullptr;
            } else {
                cs->null_count = G - v;
            }
        }
        PD_HIP(hipStreamSynchronize(s));
    } catch (...) {
        free_col(c, ck);
        free_col(c, cs);
        free_col(c, cc);
        throw;
    }
    *ok = ck;
    *os = cs;
    *oc = cc;
}

}  // namespace pd
