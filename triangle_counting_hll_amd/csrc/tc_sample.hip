// DOULION edge sampling on gfx950 (experiment/exp_gpu.py:396-450, _doulion_forward;
// src/doulion.py:11-45).
//
// The reference keeps forward edge e iff rng.random(E)[e] < q with
// rng = numpy.random.default_rng(seed): numpy's PCG64 bit generator
// (128-bit LCG, XSL-RR output) and random() = (next_uint64 >> 11) * 2^-53.
// The same stream is generated here in parallel: thread t jumps the LCG to
// draw t * kDrawsPerThread (Brown's O(log n) LCG advance) and steps from
// there, so mask[e] equals the reference's mask bit for bit.  The host reads
// the generator's initial 128-bit state and increment from numpy
// (bit_generator.state), so no seeding algorithm is restated.
//
// The sampled forward edges keep the original orientation (still acyclic),
// so the sample's triangles are counted by the exact pipeline on the
// compacted CSR; the estimate is count / q^3.
#include "tc_common.hpp"

namespace tc {

struct U128 {
  uint64_t lo, hi;
};

__device__ __forceinline__ U128 mul128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo * b.lo;
  r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
  return r;
}
__device__ __forceinline__ U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
  return r;
}

// PCG64's default 128-bit multiplier.
__device__ __forceinline__ U128 pcg_mult() {
  return U128{0x4385DF649FCCF645ull, 0x2360ED051FC65DA4ull};
}

// State after `delta` more steps of s <- s * M + inc.
__device__ __forceinline__ U128 pcg_advance(U128 state, U128 inc, uint64_t delta) {
  U128 acc_mult{1, 0}, acc_plus{0, 0};
  U128 cur_mult = pcg_mult(), cur_plus = inc;
  while (delta) {
    if (delta & 1) {
      acc_mult = mul128(acc_mult, cur_mult);
      acc_plus = add128(mul128(acc_plus, cur_mult), cur_plus);
    }
    cur_plus = mul128(add128(cur_mult, U128{1, 0}), cur_plus);
    cur_mult = mul128(cur_mult, cur_mult);
    delta >>= 1;
  }
  return add128(mul128(acc_mult, state), acc_plus);
}

__device__ __forceinline__ uint64_t pcg_output(U128 s) {
  const uint64_t x = s.hi ^ s.lo;
  const unsigned rot = (unsigned)(s.hi >> 58);
  return (x >> rot) | (x << ((64u - rot) & 63u));
}

constexpr int kDrawsPerThread = 64;

__global__ void __launch_bounds__(kBlock)
k_pcg64_below(U128 state, U128 inc, uint64_t first, int64_t n, double q,
              uint8_t *__restrict__ keep) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i0 = t * kDrawsPerThread;
  if (i0 >= n) return;
  // draw i comes from the state after first + i + 1 steps
  U128 s = pcg_advance(state, inc, first + (uint64_t)i0);
  const U128 M = pcg_mult();
  const int64_t i1 = min(n, i0 + kDrawsPerThread);
  for (int64_t i = i0; i < i1; ++i) {
    s = add128(mul128(s, M), inc);
    const double u = (double)(pcg_output(s) >> 11) * (1.0 / 9007199254740992.0);
    keep[i] = u < q ? 1 : 0;
  }
}

template <class I>
__global__ void k_sub_row_ptr(const I *__restrict__ rp, int64_t n, const int64_t *__restrict__ pos,
                              int64_t *__restrict__ out_rp) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u <= n;
       u += (int64_t)gridDim.x * blockDim.x)
    out_rp[u] = pos[(int64_t)rp[u]];
}

__global__ void k_sub_cols(const int32_t *__restrict__ col, const uint8_t *__restrict__ keep,
                           const int64_t *__restrict__ pos, int64_t E, int32_t *__restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    if (keep[e]) out[pos[e]] = col[e];
}

}  // namespace tc

extern "C" int tc_pcg64_below(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi,
                              uint64_t inc_lo, uint64_t first, int64_t n, double q,
                              uint8_t *keep, void *stream) {
  using namespace tc;
  TC_CHECK(n >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK(n == 0 || keep, TC_ERR_INVALID, "keep is required");
  TC_TRY(ensure_init());
  if (n == 0) return TC_OK;
  const int64_t threads = (n + kDrawsPerThread - 1) / kDrawsPerThread;
  hipLaunchKernelGGL(k_pcg64_below, dim3((unsigned)((threads + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, U128{state_lo, state_hi},
                     U128{inc_lo, inc_hi}, first, n, q, keep);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_edge_subgraph(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                                int64_t n_nodes, int64_t n_edges, const uint8_t *keep,
                                int64_t *out_row_ptr, int32_t *out_col, int64_t *n_kept,
                                void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(row_ptr && out_row_ptr && n_kept, TC_ERR_INVALID,
           "row_ptr, out_row_ptr and n_kept are required");
  TC_CHECK(n_nodes >= 0 && n_edges >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK(n_edges == 0 || (col_idx && keep && out_col), TC_ERR_INVALID,
           "col_idx, keep and out_col are required");
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  Scratch pos;
  TC_TRY(pos.alloc(sizeof(int64_t) * (n_edges + 1), s));
  TC_TRY(exclusive_scan<uint8_t>(keep, pos.as<int64_t>(), n_edges, nullptr, s));
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    hipLaunchKernelGGL(k_sub_row_ptr<I>, dim3(grid_for(n_nodes + 1)), dim3(kBlock), 0, s,
                       (const I *)row_ptr, n_nodes, (const int64_t *)pos.as<int64_t>(),
                       out_row_ptr);
  });
  TC_HIP(hipGetLastError());
  if (n_edges > 0) {
    hipLaunchKernelGGL(k_sub_cols, dim3(grid_for(n_edges)), dim3(kBlock), 0, s, col_idx, keep,
                       (const int64_t *)pos.as<int64_t>(), n_edges, out_col);
    TC_HIP(hipGetLastError());
  }
  if (hipMemcpyAsync(n_kept, pos.as<int64_t>() + n_edges, sizeof(int64_t), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "edge subgraph readback failed");
    return TC_ERR_HIP;
  }
  return TC_OK;
}
