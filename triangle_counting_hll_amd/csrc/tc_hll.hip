// (c) HLL sketch build, (d) FP64 union estimation, (e) hybrid exact/HLL router.
//
// Register semantics follow the reference's integer arithmetic exactly:
//   h = splitmix64 finaliser(id); j = h >> (64 - p); w = low 64 - p bits;
//   rho = (64 - p) - bit_length(w) + 1, or 65 - p when w == 0;
//   reg[j] = max(reg[j], rho)                       (src/hll.py:81-91)
// The estimator is the FP64 one of src/hll.py:37-54; its sum Z = sum_j 2^-r_j
// is accumulated in FP64 in a fixed order (deterministic, ~1e-14 relative).
//
// TC_HLL_CPU_EXP selects the CPU experiment's variant instead
// (cpu_exp/triangle_hll_experiment.py:347-398, 615-635): h = seeded splitmix64
// (id + seed + 0x9E3779B97F4A7C15, then the finaliser); j = low p bits;
// rho from the high 64 - p bits; estimator with early returns and the 2^32
// large-range constant.
//
// TC_HLL_PROTO32 selects the prototype GPU experiment's variant
// (experiment1_gpu/gpu_experiment_fixed.py:131-166, 209-254): a 32-bit mixer
// of id ^ 0x9747B28C; j = top p bits; rho = 1 + the zeros above the first set
// bit of the low 32 - p bits of (h << p) | 1; estimator with LinearCounting
// and no large-range correction; HLL support est(u) + est(v) - est(u | v),
// raised to 0 when negative (no upper clamp).
#include <algorithm>

#include "tc_common.hpp"
#include "tc_exact.hpp"
#include "tc_hll.hpp"
#include "tc_router.hpp"

namespace tc {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint32_t hll_rank(uint64_t h, int p) {
  const uint64_t w = h & ((~0ull) >> p);
  return w == 0 ? (uint32_t)(65 - p) : (uint32_t)(__clzll((long long)w) - p + 1);
}

__device__ __forceinline__ uint32_t mix32_proto(uint32_t v) {
  uint32_t h = 0x9747B28Cu ^ v;
  h = (h ^ (h >> 16)) * 0x45D9F3Bu;
  h = (h ^ (h >> 16)) * 0x45D9F3Bu;
  return h ^ (h >> 16);
}

__host__ __device__ constexpr bool valid_hash_mode(int mode) {
  return mode == TC_HLL_GPU || mode == TC_HLL_CPU_EXP || mode == TC_HLL_PROTO32;
}

// Register index and rank of one id under each hash mode.
__device__ __forceinline__ void hll_slot(int64_t id, int p, int mode, uint64_t seed, uint32_t &j,
                                         uint32_t &rho) {
  if (mode == TC_HLL_PROTO32) {
    const uint32_t h = mix32_proto((uint32_t)id);
    j = h >> (32 - p);
    const uint32_t window = ((h << p) | 1u) & ((1u << (32 - p)) - 1u);  // never zero
    rho = (uint32_t)__clz((int)window) - (uint32_t)p + 1u;
  } else if (mode == TC_HLL_CPU_EXP) {
    const uint64_t h = mix64((uint64_t)id + seed + 0x9E3779B97F4A7C15ull);
    j = (uint32_t)(h & ((1ull << p) - 1));
    const uint64_t w = h >> p;  // the high 64 - p bits
    rho = w == 0 ? (uint32_t)(65 - p) : (uint32_t)(__clzll((long long)w) - p + 1);
  } else {
    const uint64_t h = mix64((uint64_t)id);
    j = (uint32_t)(h >> (64 - p));
    rho = hll_rank(h, p);
  }
}

__global__ void k_hash64(const uint64_t *__restrict__ x, uint64_t *__restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = mix64(x[i]);
}

template <class I>
__global__ void k_sketch_flags(Rows<I> R, int64_t n, int64_t theta, uint8_t *__restrict__ has) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * blockDim.x)
    has[u] = R.deg(u) >= theta ? 1 : 0;
}

// One workgroup per sketched vertex: registers packed 4 per 32-bit LDS word,
// byte max through compare-and-swap, then one coalesced write of the row.
template <class I>
__global__ void __launch_bounds__(kBlock)
k_hll_build(Rows<I> R, const int32_t *__restrict__ col, const int32_t *__restrict__ list,
            const int64_t *__restrict__ n_list_p, int p, uint8_t *__restrict__ regs,
            const int32_t *__restrict__ slot, int accumulate, int mode, uint64_t seed) {
  extern __shared__ uint32_t t32[];
  const int m = 1 << p;
  const int words = m >> 2;
  const int64_t n_list = *n_list_p;
  for (int64_t li = blockIdx.x; li < n_list; li += gridDim.x) {
    const int32_t u = list[li];
    const int64_t r = slot ? (int64_t)slot[u] : (int64_t)u;
    if (r < 0) continue;
    uint32_t *row = (uint32_t *)(regs + r * (int64_t)m);
    for (int i = threadIdx.x; i < words; i += blockDim.x) t32[i] = accumulate ? row[i] : 0u;
    __syncthreads();
    const int64_t b = R.begin(u), e = R.end(u);
    for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      uint32_t j, rho;
      hll_slot(col[k], p, mode, seed, j, rho);
      uint32_t *wp = &t32[j >> 2];
      const int sh = (j & 3) * 8;
      uint32_t old = *wp;
      while (((old >> sh) & 0xffu) < rho) {
        const uint32_t nw = (old & ~(0xffu << sh)) | (rho << sh);
        const uint32_t prev = atomicCAS(wp, old, nw);
        if (prev == old) break;
        old = prev;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < words; i += blockDim.x) row[i] = t32[i];
    __syncthreads();
  }
}

template <int G>
__global__ void __launch_bounds__(kBlock)
k_union_estimate(const uint8_t *__restrict__ regs, const int32_t *__restrict__ slot, int p,
                 const int32_t *__restrict__ us, const int32_t *__restrict__ vs, int64_t k,
                 double *__restrict__ out, int mode) {
  const int m = 1 << p;
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t base = wid * kWave; base < k; base += nw * kWave) {
    const int64_t i = base + lane;
    const int n_valid = (int)min<int64_t>(kWave, k - base);
    int32_t a = 0, b = 0;
    if (i < k) {
      a = slot ? slot[us[i]] : us[i];
      b = slot ? slot[vs[i]] : vs[i];
    }
    const RegStats st = batch_stats<G, false>(regs, m, m, a, b, n_valid);
    if (i < k) out[i] = hll_estimate(m, st.zeros[2], st.z[2], mode);
  }
}

// ---------------------------------------------------------------------------
// Router features, candidates in batches of 64 per wavefront (batch_stats):
//   f0 = log2(1 + deg u), f1 = log2(1 + deg v), f2 = f0 - f1,
//   f3 = log2(1 + card u), f4 = log2(1 + card v)  (per-vertex estimates),
//   f5 = log2(1 + sampled union): the estimator on the first 2^s registers of
//        both rows merged -- the elements whose register index is below 2^s,
//        i.e. whose top p - s index bits are zero, a 2^-(p-s) sample of the
//        union -- scaled by 2^(p-s).  It reads 2 * 2^s bytes per edge instead
//        of 2 * 2^p.
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(kBlock)
k_router_features(const int32_t *__restrict__ degrees, const int32_t *__restrict__ eu,
                  const int32_t *__restrict__ ev, const int32_t *__restrict__ cand, int64_t n_cand,
                  const uint8_t *__restrict__ regs, const int32_t *__restrict__ slot, int p,
                  int sample_bits, const float *__restrict__ card, float *__restrict__ feats) {
  const int m = 1 << p;
  const int ms = 1 << sample_bits;
  const double scale = (double)(1 << (p - sample_bits));
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t base = wid * kWave; base < n_cand; base += nw * kWave) {
    const int64_t i = base + lane;
    const bool valid = i < n_cand;
    const int n_valid = (int)min<int64_t>(kWave, n_cand - base);
    int32_t u = 0, v = 0, a = 0, b = 0;
    if (valid) {
      const int32_t e = cand[i];
      u = eu[e];
      v = ev[e];
      a = slot ? slot[u] : u;
      b = slot ? slot[v] : v;
    }
    const RegStats st = batch_stats<G, false>(regs, m, ms, a, b, n_valid);
    if (!valid) continue;
    const double us = hll_estimate(ms, st.zeros[2], st.z[2]) * scale;
    float f[kIn];
    router_features(degrees[u], degrees[v], card[u], card[v], us, f);
#pragma unroll
    for (int q = 0; q < kIn; ++q) feats[i * kIn + q] = f[q];
  }
}

// ---------------------------------------------------------------------------
// Hybrid
// ---------------------------------------------------------------------------

// Degree used by the router and the support formula: the caller's `degrees`
// array when given (the reference passes forward degrees), else the row length.
template <class I> struct Degrees {
  Rows<I> R;
  const int32_t *d;
  __device__ __forceinline__ int64_t operator()(int64_t u) const {
    return d ? (int64_t)d[u] : R.deg(u);
  }
};

// Routing flags: use[e] = wants HLL && both sketched.  counters[0] +=
// HLL-routed edges, counters[1] += fallbacks (wanted HLL, missing sketch).
template <class I>
__global__ void k_route(Degrees<I> D, const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                        int64_t E, const uint8_t *__restrict__ has, int64_t gamma,
                        const uint8_t *__restrict__ route, uint8_t *__restrict__ use,
                        unsigned long long *__restrict__ counters) {
  uint64_t n_use = 0, n_fb = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = eu[e], v = ev[e];
    // min(D(u), D(v)) >= gamma, reading D(v) only once D(u) reaches gamma and
    // the sketch flags only for edges that want HLL
    bool wants;
    if (route) {
      wants = route[e] != 0;
    } else {
      wants = D(u) >= gamma && D(v) >= gamma;
    }
    const bool both = wants && has[u] == 1 && has[v] == 1;
    const bool hll = wants && both;
    use[e] = hll ? 1 : 0;
    n_use += hll;
    n_fb += (wants && !both);
  }
  n_use = wave_sum(n_use);
  n_fb = wave_sum(n_fb);
  if (lane_id() == 0) {
    if (n_use) atomicAdd(&counters[0], (unsigned long long)n_use);
    if (n_fb) atomicAdd(&counters[1], (unsigned long long)n_fb);
  }
}

// HLL-routed edges in batches of 64 per wavefront (batch_stats): support
// estimate, clamp, clamp count.  kTriple: the prototype's support from three
// estimates (TC_HLL_PROTO32).
template <class I, int G, bool kTriple>
__global__ void __launch_bounds__(kBlock)
k_hll_edges(Degrees<I> D, const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
            const int32_t *__restrict__ list, const int64_t *__restrict__ n_list_p,
            const uint8_t *__restrict__ regs, const int32_t *__restrict__ slot, int p,
            double *__restrict__ s_out, double *__restrict__ per_edge,
            unsigned long long *__restrict__ n_clamped, int mode) {
  const int m = 1 << p;
  const int lane = lane_id();
  const int64_t n_list = *n_list_p;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  uint64_t clamped = 0;
  for (int64_t base = wid * kWave; base < n_list; base += nw * kWave) {
    const int64_t i = base + lane;
    const bool valid = i < n_list;
    const int n_valid = (int)min<int64_t>(kWave, n_list - base);
    int32_t e = 0, u = 0, v = 0, a = 0, b = 0;
    if (valid) {
      e = list[i];
      u = eu[e];
      v = ev[e];
      a = slot ? slot[u] : u;
      b = slot ? slot[v] : v;
    }
    const RegStats st = batch_stats<G, kTriple>(regs, m, m, a, b, n_valid);
    if (!valid) continue;
    double s;
    if (kTriple) {
      s = hll_estimate(m, st.zeros[0], st.z[0], mode) + hll_estimate(m, st.zeros[1], st.z[1], mode) -
          hll_estimate(m, st.zeros[2], st.z[2], mode);
      if (s < 0.0) {
        s = 0.0;
        ++clamped;
      }
    } else {
      const double E = hll_estimate(m, st.zeros[2], st.z[2], mode);
      const int64_t du = D(u), dv = D(v);
      s = (double)du + (double)dv - E;
      const double upper = (double)min(du, dv);
      if (s < 0.0) {
        s = 0.0;
        ++clamped;
      } else if (s > upper) {
        s = upper;
        ++clamped;
      }
    }
    s_out[i] = s;
    if (per_edge) per_edge[e] = s;
  }
  clamped = wave_sum(clamped);
  if (lane == 0 && clamped) atomicAdd(n_clamped, (unsigned long long)clamped);
}

template <class I>
__global__ void k_exact_per_edge(Rows<I> R, const int32_t *__restrict__ col,
                                 const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                                 int64_t E, const uint8_t *__restrict__ use,
                                 double *__restrict__ per_edge) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (use && use[e]) continue;
    int64_t a0 = R.begin(eu[e]), a1 = R.end(eu[e]), b0 = R.begin(ev[e]), b1 = R.end(ev[e]);
    if (a1 - a0 > b1 - b0) {
      int64_t t0 = a0, t1 = a1;
      a0 = b0; a1 = b1; b0 = t0; b1 = t1;
    }
    int c = 0;
    for (int64_t i = a0; i < a1; ++i) {
      const int x = col[i];
      int64_t lo = b0, hi = b1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < x) lo = mid + 1; else hi = mid;
      }
      if (lo < b1 && col[lo] == x) ++c;
    }
    per_edge[e] = (double)c;
  }
}

// Deterministic FP64 sum: fixed-size tiles summed in a fixed tree order, then
// one workgroup folds the tile sums with Neumaier compensation.
constexpr int kSumTile = 2048;

__device__ __forceinline__ double block_sum_fixed(double v) {
  __shared__ double buf[kBlock];
  buf[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) buf[threadIdx.x] += buf[threadIdx.x + s];
    __syncthreads();
  }
  double r = buf[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kBlock)
k_tile_sums_f64(const double *__restrict__ x, const int64_t *__restrict__ n_p,
                double *__restrict__ partial, int64_t max_tiles) {
  const int64_t n = *n_p;
  const int64_t tiles = (n + kSumTile - 1) / kSumTile;
  for (int64_t t = blockIdx.x; t < tiles && t < max_tiles; t += gridDim.x) {
    double v = 0.0;
    for (int64_t i = t * kSumTile + threadIdx.x; i < min(n, (t + 1) * kSumTile); i += kBlock)
      v += x[i];
    v = block_sum_fixed(v);
    if (threadIdx.x == 0) partial[t] = v;
  }
}

__global__ void __launch_bounds__(kBlock)
k_hybrid_finish(const double *__restrict__ partial, const int64_t *__restrict__ n_hll_p,
                const unsigned long long *__restrict__ exact_sum,
                const unsigned long long *__restrict__ counters, int64_t E, int part,
                double *__restrict__ counts, int64_t *__restrict__ stats,
                int64_t *__restrict__ fallback) {
  const int64_t n = *n_hll_p;
  const int64_t tiles = (n + kSumTile - 1) / kSumTile;
  double s = 0.0, c = 0.0;  // Neumaier
  for (int64_t t = threadIdx.x; t < tiles; t += kBlock) {
    const double x = partial[t];
    const double y = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - y) + x : (x - y) + s;
    s = y;
  }
  __shared__ double bs[kBlock], bc[kBlock];
  bs[threadIdx.x] = s;
  bc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int i = 0; i < kBlock; ++i) {
      const double x = bs[i];
      const double y = S + x;
      C += (fabs(S) >= fabs(x)) ? (S - y) + x : (x - y) + S;
      S = y;
      C += bc[i];
    }
    const double hll_sum = S + C;
    const unsigned long long n_hll = counters[0];
    double add = 0.0;
    if (part != TC_PART_HLL) add += (double)exact_sum[0];
    if (part != TC_PART_EXACT) add += hll_sum;
    counts[0] += add;
    if (part != TC_PART_EXACT) {
      stats[0] += (int64_t)n_hll;
      stats[2] += (int64_t)counters[2];
    }
    if (part != TC_PART_HLL) stats[1] += E - (int64_t)n_hll;
    if (fallback) fallback[0] += (int64_t)counters[1];
  }
}


}  // namespace tc

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" int tc_hash64(const uint64_t *x, uint64_t *out, int64_t n, void *stream) {
  using namespace tc;
  TC_CHECK(n >= 0 && (n == 0 || (x && out)), TC_ERR_INVALID, "invalid arguments to tc_hash64");
  TC_TRY(ensure_init());
  if (n == 0) return TC_OK;
  hipLaunchKernelGGL(k_hash64, dim3(grid_for(n, kBlock, 65536)), dim3(kBlock), 0,
                     (hipStream_t)stream, x, out, n);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_hll_build(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                            int64_t n_nodes, int64_t theta, int p, uint8_t *regs,
                            const int32_t *slot, uint8_t *has_sketch, int accumulate,
                            int hash_mode, uint64_t seed, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(p >= 4 && p <= 16, TC_ERR_INVALID, "p must be in [4, 16], got %d", p);
  TC_CHECK(valid_hash_mode(hash_mode), TC_ERR_INVALID, "unknown hash mode %d", hash_mode);
  TC_CHECK(row_ptr && has_sketch && n_nodes >= 0, TC_ERR_INVALID,
           "row_ptr and has_sketch are required");
  TC_CHECK(n_nodes == 0 || regs, TC_ERR_INVALID, "regs is required");
  TC_CHECK(((uintptr_t)regs & 3) == 0, TC_ERR_INVALID, "regs must be 4-byte aligned");
  TC_TRY(ensure_init());
  if (n_nodes == 0) return TC_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch list, cnt;
  TC_TRY(list.alloc(sizeof(int32_t) * n_nodes, s));
  TC_TRY(cnt.alloc(sizeof(int64_t), s));
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    Rows<I> R{(const I *)row_ptr};
    hipLaunchKernelGGL(k_sketch_flags<I>, dim3(grid_for(n_nodes, kBlock, 16384)), dim3(kBlock),
                       0, s, R, n_nodes, theta, has_sketch);
    TC_HIP(hipGetLastError());
    TC_TRY(compact_indices(has_sketch, n_nodes, list.as<int32_t>(), cnt.as<int64_t>(), s));
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(n_nodes, num_cus() * 8));
    {
      StageTimer _timer(TC_STAGE_HLL_BUILD, s);
      hipLaunchKernelGGL(k_hll_build<I>, dim3(g), dim3(kBlock), (size_t)1 << p, s, R, col_idx,
                         (const int32_t *)list.as<int32_t>(), (const int64_t *)cnt.as<int64_t>(), p,
                         regs, slot, accumulate, hash_mode, seed);
    }
    TC_HIP(hipGetLastError());
  });
  return TC_OK;
}

extern "C" int tc_union_estimate(const uint8_t *regs, const int32_t *slot, int p,
                                 const int32_t *us, const int32_t *vs, int64_t k, double *out,
                                 int hash_mode, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(p >= 4 && p <= 16, TC_ERR_INVALID, "p must be in [4, 16], got %d", p);
  TC_CHECK(valid_hash_mode(hash_mode), TC_ERR_INVALID, "unknown hash mode %d", hash_mode);
  TC_CHECK(k >= 0 && (k == 0 || (regs && us && vs && out)), TC_ERR_INVALID,
           "invalid arguments to tc_union_estimate");
  TC_CHECK(((uintptr_t)regs & 15) == 0, TC_ERR_INVALID, "regs must be 16-byte aligned");
  TC_TRY(ensure_init());
  if (k == 0) return TC_OK;
  const unsigned g = (unsigned)std::min<int64_t>((k + kBlock - 1) / kBlock, num_cus() * 16);
  TC_DISPATCH_GROUP(p, G, {
    hipLaunchKernelGGL(k_union_estimate<G>, dim3(g), dim3(kBlock), 0, (hipStream_t)stream, regs,
                       slot, p, us, vs, k, out, hash_mode);
  });
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_router_features(const int32_t *degrees, const int32_t *edge_u,
                                  const int32_t *edge_v, const int32_t *cand, int64_t n_cand,
                                  const uint8_t *regs, const int32_t *slot, int p, int sample_bits,
                                  const float *card, float *feats, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(p >= 4 && p <= 16 && sample_bits >= 4 && sample_bits <= p, TC_ERR_INVALID,
           "need 4 <= sample_bits <= p <= 16 (got p=%d, sample_bits=%d)", p, sample_bits);
  TC_CHECK(n_cand >= 0 && (n_cand == 0 || (degrees && edge_u && edge_v && cand && regs && card &&
                                           feats)),
           TC_ERR_INVALID, "invalid arguments to tc_router_features");
  TC_CHECK(((uintptr_t)regs & 15) == 0, TC_ERR_INVALID, "regs must be 16-byte aligned");
  TC_TRY(ensure_init());
  if (n_cand == 0) return TC_OK;
  const unsigned g =
      (unsigned)std::min<int64_t>((n_cand + kBlock - 1) / kBlock, num_cus() * 16);
  TC_DISPATCH_GROUP(sample_bits, G, {
    hipLaunchKernelGGL(k_router_features<G>, dim3(g), dim3(kBlock), 0, (hipStream_t)stream,
                       degrees, edge_u, edge_v, cand, n_cand, regs, slot, p, sample_bits, card,
                       feats);
  });
  TC_HIP(hipGetLastError());
  return TC_OK;
}

namespace tc {

// target[0] += source[0] (a device counter handed to the caller).
__global__ void k_add_counter(const unsigned long long *__restrict__ source,
                              int64_t *__restrict__ target) {
  target[0] += (int64_t)source[0];
}

// tc_hybrid_count, and with net (net->params != null) the learned router's
// count: routing by the network inside the exact pass (part must be BOTH).
// use_out (uint8 [E], optional) receives the per-edge route taken; n_cand
// (device int64, optional) += the network's candidate edges.
static int hybrid_impl(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                       int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                       int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                       const int32_t *slot, const uint8_t *has_sketch, int64_t gamma, int p,
                       const uint8_t *route, int part, double *counts, int64_t *stats,
                       int64_t *fallback, double *per_edge, int hash_mode, hipStream_t s,
                       const RouterNet *net, uint8_t *use_out, int64_t *n_cand,
                       const tc_exact_params *params = nullptr) {
  TC_CHECK(valid_hash_mode(hash_mode), TC_ERR_INVALID, "unknown hash mode %d", hash_mode);
  TC_CHECK(part == TC_PART_BOTH || part == TC_PART_EXACT || part == TC_PART_HLL, TC_ERR_INVALID,
           "unknown part %d", part);
  TC_CHECK(part == TC_PART_EXACT || (p >= 4 && p <= 16), TC_ERR_INVALID,
           "p must be in [4, 16], got %d", p);
  TC_CHECK(counts && stats && has_sketch && n_edges >= 0 && n_nodes >= 0, TC_ERR_INVALID,
           "invalid arguments to tc_hybrid_count");
  TC_CHECK(n_edges == 0 || (edge_u && edge_v), TC_ERR_INVALID, "edge_u and edge_v are required");
  TC_CHECK(n_edges == 0 || n_nodes > 0, TC_ERR_INVALID, "edges need n_nodes > 0");
  TC_CHECK(part == TC_PART_HLL || n_edges == 0 || (row_ptr && col_idx), TC_ERR_INVALID,
           "row_ptr and col_idx are required for the exact part");
  TC_CHECK(row_ptr || degrees, TC_ERR_INVALID, "row_ptr or degrees is required");
  TC_CHECK(part == TC_PART_EXACT || n_edges == 0 || regs, TC_ERR_INVALID,
           "sketch registers are required for the HLL part");
  TC_CHECK(((uintptr_t)regs & 15) == 0, TC_ERR_INVALID, "regs must be 16-byte aligned");
  TC_CHECK(n_edges < INT32_MAX, TC_ERR_INVALID, "edge ids must fit int32");
  TC_TRY(ensure_init());
  const int64_t E = n_edges;
  Scratch use_buf, list, nlist, svals, partial, acc;
  uint8_t *use = use_out;
  if (!use) {
    TC_TRY(use_buf.alloc(std::max<int64_t>(E, 1), s));
    use = use_buf.as<uint8_t>();
  }
  TC_TRY(list.alloc(sizeof(int32_t) * std::max<int64_t>(E, 1), s));
  TC_TRY(nlist.alloc(sizeof(int64_t), s));
  TC_TRY(svals.alloc(sizeof(double) * std::max<int64_t>(E, 1), s));
  const int64_t max_tiles = (E + kSumTile - 1) / kSumTile + 1;
  TC_TRY(partial.alloc(sizeof(double) * max_tiles, s));
  TC_TRY(acc.alloc(sizeof(unsigned long long) * 5, s));
  unsigned long long *exact_sum = acc.as<unsigned long long>();
  // [n_hll, n_fallback, n_clamped, n_candidates]
  unsigned long long *counters = exact_sum + 1;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, exact_sum, 5, 0ull);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int32_t *)nlist.ptr, 2, 0);
  TC_HIP(hipGetLastError());
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    Rows<I> R{(const I *)row_ptr};
    Degrees<I> D{R, degrees};
    if (E > 0) {
      if (part == TC_PART_HLL) {
        hipLaunchKernelGGL(k_route<I>, dim3(grid_for(E, kBlock, 16384)), dim3(kBlock), 0, s, D,
                           edge_u, edge_v, E, has_sketch, gamma, route, use, counters);
        TC_HIP(hipGetLastError());
      } else {
        // the exact pipeline's edge pass routes each edge (same rule as
        // k_route) and counts only the edges it keeps
        HybridRoute hr{degrees, has_sketch, gamma, route, use, counters, RouterNet{}};
        if (net) hr.net = *net;
        TC_TRY(exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, E,
                                nullptr, nullptr, exact_sum, params, s, &hr));
        if (per_edge)
          hipLaunchKernelGGL(k_exact_per_edge<I>, dim3(grid_for(E, kBlock, 16384)), dim3(kBlock),
                             0, s, R, col_idx, edge_u, edge_v, E,
                             (const uint8_t *)use, per_edge);
      }
      if (part != TC_PART_EXACT) {
        TC_TRY(compact_indices(use, E, list.as<int32_t>(), nlist.as<int64_t>(), s));
        // the HLL edge count is on the device: size the grid for all E edges
        const unsigned gw = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>((E + kBlock - 1) / kBlock, num_cus() * 16));
        {
          StageTimer _timer(TC_STAGE_HLL_EDGES, s);
          const int32_t *lp = list.as<int32_t>();
          const int64_t *np_ = nlist.as<int64_t>();
          if (hash_mode == TC_HLL_PROTO32) {
            TC_DISPATCH_GROUP(p, G, {
              hipLaunchKernelGGL((k_hll_edges<I, G, true>), dim3(gw), dim3(kBlock), 0, s, D, edge_u,
                                 edge_v, lp, np_, regs, slot, p, svals.as<double>(), per_edge,
                                 counters + 2, hash_mode);
            });
          } else {
            TC_DISPATCH_GROUP(p, G, {
              hipLaunchKernelGGL((k_hll_edges<I, G, false>), dim3(gw), dim3(kBlock), 0, s, D,
                                 edge_u, edge_v, lp, np_, regs, slot, p, svals.as<double>(),
                                 per_edge, counters + 2, hash_mode);
            });
          }
        }
        hipLaunchKernelGGL(k_tile_sums_f64, dim3((unsigned)std::min<int64_t>(max_tiles, 65536)),
                           dim3(kBlock), 0, s, (const double *)svals.as<double>(),
                           (const int64_t *)nlist.as<int64_t>(), partial.as<double>(), max_tiles);
        TC_HIP(hipGetLastError());
      }
    }
  });
  hipLaunchKernelGGL(k_hybrid_finish, dim3(1), dim3(kBlock), 0, s,
                     (const double *)partial.as<double>(), (const int64_t *)nlist.as<int64_t>(),
                     (const unsigned long long *)exact_sum, (const unsigned long long *)counters, E,
                     part, counts, stats, fallback);
  TC_HIP(hipGetLastError());
  if (n_cand)
    hipLaunchKernelGGL(k_add_counter, dim3(1), dim3(1), 0, s,
                       (const unsigned long long *)(counters + 3), n_cand);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

}  // namespace tc

extern "C" int tc_hybrid_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                               int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                               int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                               const int32_t *slot, const uint8_t *has_sketch, int64_t gamma, int p,
                               const uint8_t *route, int part, double *counts, int64_t *stats,
                               int64_t *fallback, double *per_edge, int hash_mode,
                               void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  return hybrid_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges, degrees,
                     regs, slot, has_sketch, gamma, p, route, part, counts, stats, fallback,
                     per_edge, hash_mode, (hipStream_t)stream, nullptr, nullptr, nullptr);
}

extern "C" int tc_hybrid_count_tuned(const void *row_ptr, int row_ptr_64,
                                     const int32_t *col_idx, int64_t n_nodes,
                                     const int32_t *edge_u, const int32_t *edge_v,
                                     int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                                     const int32_t *slot, const uint8_t *has_sketch,
                                     int64_t gamma, int p, const uint8_t *route, int part,
                                     double *counts, int64_t *stats, int64_t *fallback,
                                     double *per_edge, int hash_mode,
                                     const tc_exact_params *params, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  return hybrid_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges, degrees,
                     regs, slot, has_sketch, gamma, p, route, part, counts, stats, fallback,
                     per_edge, hash_mode, (hipStream_t)stream, nullptr, nullptr, nullptr, params);
}

extern "C" int tc_router_hybrid_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                                      int64_t n_nodes, const int32_t *edge_u,
                                      const int32_t *edge_v, int64_t n_edges,
                                      const int32_t *degrees, const uint8_t *regs,
                                      const int32_t *slot, const uint8_t *has_sketch, int p,
                                      const tc_router_net *net, double *counts, int64_t *stats,
                                      int64_t *n_candidates, uint8_t *use, double *per_edge,
                                      void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(net && net->params && net->norm && net->card, TC_ERR_INVALID,
           "the router network needs params, norm and card");
  TC_CHECK(p >= 4 && p <= 16 && net->sample_bits >= 4 && net->sample_bits <= p, TC_ERR_INVALID,
           "need 4 <= sample_bits <= p <= 16 (got p=%d, sample_bits=%d)", p, net->sample_bits);
  TC_CHECK(net->min_cost_degree >= 0, TC_ERR_INVALID, "min_cost_degree must be >= 0");
  TC_CHECK(n_edges == 0 || (row_ptr && col_idx && edge_u && edge_v && regs), TC_ERR_INVALID,
           "row_ptr, col_idx, edge_u, edge_v and regs are required");
  const RouterNet rn{net->params, net->norm, net->card, regs, slot, p, net->sample_bits,
                     net->threshold, net->min_cost_degree};
  return hybrid_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges, degrees,
                     regs, slot, has_sketch, 0, p, nullptr, TC_PART_BOTH, counts, stats, nullptr,
                     per_edge, TC_HLL_GPU, (hipStream_t)stream, &rn, use, n_candidates);
}
