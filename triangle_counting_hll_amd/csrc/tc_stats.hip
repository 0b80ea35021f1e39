// Degree statistics of an undirected edge list on gfx950
// (experiment/gpu_graph_stats.py:89-140: degree_count_kernel,
// degree_stats_kernel, degree_hist_kernel).
//
// Degree counting is one atomic per endpoint, as in the reference.  The
// summary and the histograms are reduced per workgroup first (wavefront
// shuffles, LDS) and then merged with one global atomic per workgroup and
// statistic, or per non-empty LDS bin, instead of one global atomic per vertex.
#include "tc_common.hpp"
#include "tc_exact.hpp"

namespace tc {

constexpr int kLdsBins = 8192;  // degrees below this are histogrammed in LDS
constexpr int kMaxLogBuckets = 64;

__global__ void k_degree_count(const int32_t *__restrict__ src, const int32_t *__restrict__ dst,
                               int64_t stride, int64_t m, int32_t *__restrict__ deg,
                               int64_t n_slots, unsigned int *__restrict__ bad) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = src[e * stride], v = dst[e * stride];
    if (u < 0 || v < 0 || u >= n_slots || v >= n_slots) {
      atomicOr(bad, 1u);
      continue;
    }
    atomicAdd(&deg[u], 1);
    atomicAdd(&deg[v], 1);
  }
}

__global__ void k_init_summary(unsigned long long *sums, int32_t *minmax) {
  if (threadIdx.x < 3) sums[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    minmax[0] = 0;
    minmax[1] = 0x7fffffff;
  }
}

__global__ void __launch_bounds__(kBlock)
k_degree_summary(const int32_t *__restrict__ deg, int64_t n, unsigned long long *__restrict__ sums,
                 int32_t *__restrict__ minmax) {
  __shared__ unsigned long long s_sum[3][kBlock / kWave];
  __shared__ int s_mm[2][kBlock / kWave];
  uint64_t active = 0, sum = 0, sumsq = 0;
  int dmax = 0, dmin = 0x7fffffff;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int d = deg[i];
    if (d <= 0) continue;
    active += 1;
    sum += (uint64_t)d;
    sumsq += (uint64_t)d * (uint64_t)d;
    dmax = max(dmax, d);
    dmin = min(dmin, d);
  }
  active = wave_sum(active);
  sum = wave_sum(sum);
  sumsq = wave_sum(sumsq);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dmax = max(dmax, __shfl_xor(dmax, off, kWave));
    dmin = min(dmin, __shfl_xor(dmin, off, kWave));
  }
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_sum[0][w] = active;
    s_sum[1][w] = sum;
    s_sum[2][w] = sumsq;
    s_mm[0][w] = dmax;
    s_mm[1][w] = dmin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / kWave; ++k) {
      active += s_sum[0][k];
      sum += s_sum[1][k];
      sumsq += s_sum[2][k];
      dmax = max(dmax, s_mm[0][k]);
      dmin = min(dmin, s_mm[1][k]);
    }
    if (active) {
      atomicAdd(&sums[0], (unsigned long long)active);
      atomicAdd(&sums[1], (unsigned long long)sum);
      atomicAdd(&sums[2], (unsigned long long)sumsq);
      atomicMax(&minmax[0], dmax);
      atomicMin(&minmax[1], dmin);
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_degree_hist(const int32_t *__restrict__ deg, int64_t n, unsigned long long *__restrict__ hist,
              int64_t n_hist, unsigned long long *__restrict__ log_hist, int n_log,
              unsigned int *__restrict__ bad) {
  __shared__ int s_bins[kLdsBins];
  __shared__ int s_log[kMaxLogBuckets];
  for (int i = threadIdx.x; i < kLdsBins; i += blockDim.x) s_bins[i] = 0;
  if (threadIdx.x < kMaxLogBuckets) s_log[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int d = deg[i];
    if (d <= 0) continue;
    if (d >= n_hist) {
      atomicOr(bad, 1u);
      continue;
    }
    if (d < kLdsBins)
      atomicAdd(&s_bins[d], 1);
    else
      atomicAdd(&hist[d], 1ull);
    const int b = min(31 - __clz(d), n_log - 1);  // floor(log2 d), clamped
    atomicAdd(&s_log[b], 1);
  }
  __syncthreads();
  const int lim = (int)min<int64_t>(kLdsBins, n_hist);
  for (int i = threadIdx.x; i < lim; i += blockDim.x)
    if (s_bins[i]) atomicAdd(&hist[i], (unsigned long long)s_bins[i]);
  if (threadIdx.x < n_log && s_log[threadIdx.x])
    atomicAdd(&log_hist[threadIdx.x], (unsigned long long)s_log[threadIdx.x]);
}

static int read_flag(unsigned int *d_flag, hipStream_t s, unsigned int *out) {
  if (hipMemcpyAsync(out, d_flag, sizeof(unsigned int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "flag readback failed: %s", hipGetErrorString(hipGetLastError()));
    return TC_ERR_HIP;
  }
  return TC_OK;
}

}  // namespace tc

extern "C" int tc_degree_count(const int32_t *src, const int32_t *dst, int64_t stride,
                               int64_t n_edges, int32_t *degree, int64_t n_slots, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_edges >= 0 && n_slots >= 0 && stride >= 1, TC_ERR_INVALID, "bad size or stride");
  TC_CHECK(n_edges == 0 || (src && dst && degree), TC_ERR_INVALID,
           "src, dst and degree are required");
  TC_TRY(ensure_init());
  if (n_edges == 0) return TC_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch flag;
  TC_TRY(flag.alloc(sizeof(unsigned int), s));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flag.as<int32_t>(), 1, 0);
  hipLaunchKernelGGL(k_degree_count, dim3(grid_for(n_edges, kBlock, 65536)), dim3(kBlock), 0, s,
                     src, dst, stride, n_edges, degree, n_slots, flag.as<unsigned int>());
  TC_HIP(hipGetLastError());
  unsigned int bad = 0;
  TC_TRY(read_flag(flag.as<unsigned int>(), s, &bad));
  TC_CHECK(!bad, TC_ERR_INVALID, "edge endpoint outside [0, %lld)", (long long)n_slots);
  return TC_OK;
}

extern "C" int tc_degree_summary(const int32_t *degree, int64_t n_slots, int64_t *sums,
                                 int32_t *minmax, void *stream) {
  using namespace tc;
  TC_CHECK(n_slots >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK(sums && minmax && (n_slots == 0 || degree), TC_ERR_INVALID,
           "degree, sums and minmax are required");
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_init_summary, dim3(1), dim3(64), 0, s, (unsigned long long *)sums, minmax);
  if (n_slots > 0)
    hipLaunchKernelGGL(k_degree_summary, dim3(grid_for(n_slots, kBlock, 4 * num_cus())),
                       dim3(kBlock), 0, s, degree, n_slots, (unsigned long long *)sums, minmax);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_degree_histogram(const int32_t *degree, int64_t n_slots, int64_t *hist,
                                   int64_t n_hist, int64_t *log_hist, int n_log_buckets,
                                   void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_slots >= 0 && n_hist >= 1, TC_ERR_INVALID, "bad size");
  TC_CHECK(n_log_buckets >= 1 && n_log_buckets <= kMaxLogBuckets, TC_ERR_INVALID,
           "n_log_buckets must be in [1, %d]", kMaxLogBuckets);
  TC_CHECK(hist && log_hist && (n_slots == 0 || degree), TC_ERR_INVALID,
           "degree, hist and log_hist are required");
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  Scratch flag;
  TC_TRY(flag.alloc(sizeof(unsigned int), s));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flag.as<int32_t>(), 1, 0);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n_hist)), dim3(kBlock), 0, s,
                     (unsigned long long *)hist, n_hist, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (unsigned long long *)log_hist,
                     (int64_t)n_log_buckets, 0ull);
  if (n_slots > 0)
    hipLaunchKernelGGL(k_degree_hist, dim3(grid_for(n_slots, kBlock, 4 * num_cus())),
                       dim3(kBlock), 0, s, degree, n_slots, (unsigned long long *)hist, n_hist,
                       (unsigned long long *)log_hist, n_log_buckets, flag.as<unsigned int>());
  TC_HIP(hipGetLastError());
  unsigned int bad = 0;
  TC_TRY(read_flag(flag.as<unsigned int>(), s, &bad));
  TC_CHECK(!bad, TC_ERR_INVALID, "a degree is not below the histogram length %lld",
           (long long)n_hist);
  return TC_OK;
}
