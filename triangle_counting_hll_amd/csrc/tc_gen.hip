// Device generator of raw R-MAT / Erdos-Renyi pairs for many graphs at once.
// Same counter-based stream as the numpy generators of graphgen.py:
//   draw(key, i) = mix(key + (i + 1) * GOLDEN)   (splitmix64 sequence)
// so the device and host produce identical pairs.
#include "tc_common.hpp"

namespace tc {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t gen_mix(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_generate(const uint8_t *__restrict__ kind, const int32_t *__restrict__ param,
                           const int64_t *__restrict__ id_off, const int64_t *__restrict__ pair_off,
                           const uint64_t *__restrict__ key, int G, uint32_t t_a, uint32_t t_ab,
                           uint32_t t_abc, int32_t *__restrict__ out) {
  const int64_t total = pair_off[G];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    // graph of pair i: last g with pair_off[g] <= i
    int lo = 0, hi = G;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pair_off[mid] <= i) lo = mid; else hi = mid;
    }
    const int g = lo;
    const uint64_t local = (uint64_t)(i - pair_off[g]);
    const uint64_t k = key[g];
    uint32_t u, v;
    if (kind[g] == 0) {  // Erdos-Renyi: one draw per pair
      const uint64_t r = gen_mix(k + (local + 1) * kGolden);
      const uint64_t n = (uint64_t)param[g];
      u = (uint32_t)(((r & 0xffffffffull) * n) >> 32);
      v = (uint32_t)(((r >> 32) * n) >> 32);
    } else {  // R-MAT: two levels per draw, most significant level first
      const int scale = param[g];
      const uint64_t D = (uint64_t)((scale + 1) >> 1);
      u = 0;
      v = 0;
      for (int j = 0; j < (int)D; ++j) {
        const uint64_t r64 = gen_mix(k + (local * D + (uint64_t)j + 1) * kGolden);
        for (int half = 0; half < 2; ++half) {
          const int level = 2 * j + half;
          if (level >= scale) break;
          const uint32_t r = (uint32_t)(r64 >> (32 * half));
          const uint32_t bit = 1u << (scale - 1 - level);
          if (r >= t_ab) u |= bit;
          if ((r >= t_a && r < t_ab) || r >= t_abc) v |= bit;
        }
      }
    }
    out[2 * i] = (int32_t)(id_off[g] + u);
    out[2 * i + 1] = (int32_t)(id_off[g] + v);
  }
}

}  // namespace tc

extern "C" int tc_generate_pairs(const uint8_t *kind, const int32_t *param, const int64_t *id_off,
                                 const int64_t *pair_off, const uint64_t *key, int n_graphs,
                                 uint32_t t_a, uint32_t t_ab, uint32_t t_abc, int32_t *pairs,
                                 int64_t n_pairs, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_graphs > 0 && kind && param && id_off && pair_off && key && n_pairs >= 0 &&
               (n_pairs == 0 || pairs),
           TC_ERR_INVALID, "invalid arguments to tc_generate_pairs");
  TC_TRY(ensure_init());
  if (n_pairs == 0) return TC_OK;
  hipLaunchKernelGGL(k_generate, dim3(grid_for(n_pairs, kBlock, 65536)), dim3(kBlock), 0,
                     (hipStream_t)stream, kind, param, id_off, pair_off, key, n_graphs, t_a, t_ab,
                     t_abc, pairs);
  TC_HIP(hipGetLastError());
  return TC_OK;
}
