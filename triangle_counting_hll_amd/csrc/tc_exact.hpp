// Internal interface of the exact-count pipeline, shared by the hybrid and
// batched entry points.
#pragma once

#include "tc_common.hpp"

namespace tc {

// count[0] (or count[graph_of[center]] when graph_of != nullptr) +=
// sum over edges e with !(skip && skip[e]) of |N+(eu[e]) ∩ N+(ev[e])|.
// eu == nullptr means the CSR's own COO (ev is then ignored).
int exact_count_impl(const void *row_ptr, int row_ptr_64, const int32_t *col, int64_t n,
                     const int32_t *eu, const int32_t *ev, int64_t E, const uint8_t *skip,
                     const int32_t *graph_of, unsigned long long *count,
                     const tc_exact_params *params, hipStream_t s);

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v);
__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v);

}  // namespace tc
