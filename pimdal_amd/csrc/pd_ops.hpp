// pd_ops.hpp — internal host entry points of the operators (stream-ordered,
// throwing pd::Error); the C ABI wraps these.
#pragma once
#include "pd_runtime.hpp"

namespace pd {

void select_range(pimdal_ctx* c, pimdal_type kt, const void* key, const uint32_t* pay, uint64_t n, int64_t lo,
                  int64_t hi, uint32_t* out_pay, uint32_t* out_rid, uint64_t* d_count, hipStream_t s);

void sort_pairs(pimdal_ctx* c, const uint32_t* kin, const void* vin, uint32_t vw, uint64_t n, uint32_t* kout,
                void* vout, int begin_bit, int end_bit, hipStream_t s);

}  // namespace pd
