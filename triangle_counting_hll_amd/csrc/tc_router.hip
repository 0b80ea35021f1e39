// Learned exact-vs-sketch router: a 6 -> 64 -> 1 MLP (ReLU, sigmoid).
// FP32 training (forward + backward with deterministic reductions, Adam) and
// BF16 inference.  No MFMA: the network is 513 parameters and every sample is
// independent, so each lane evaluates one sample from parameters held in LDS.
// Parameter layout and the shared inference code: tc_router.hpp.
#include <algorithm>

#include "tc_common.hpp"
#include "tc_router.hpp"

namespace tc {

constexpr int kSlots = (kParams + kWave - 1) / kWave;  // per-lane gradient slots (9)

__device__ __forceinline__ float softplus_f(float x) {
  return x > 0.f ? x + log1pf(expf(-x)) : log1pf(expf(x));
}

__device__ __forceinline__ void load_params(const float *__restrict__ params,
                                            const float *__restrict__ norm, float *sp,
                                            float *snorm) {
  for (int k = threadIdx.x; k < kParams; k += blockDim.x) sp[k] = params[k];
  for (int k = threadIdx.x; k < 2 * kIn; k += blockDim.x) snorm[k] = norm[k];
  __syncthreads();
}

__device__ __forceinline__ void normalise(const float *__restrict__ feats, int64_t s,
                                          const float *snorm, float x[kIn]) {
#pragma unroll
  for (int i = 0; i < kIn; ++i) x[i] = (feats[s * kIn + i] - snorm[i]) * snorm[kIn + i];
}

__global__ void __launch_bounds__(kBlock)
k_mlp_forward(const float *__restrict__ params, const float *__restrict__ norm,
              const float *__restrict__ feats, int64_t n, float *__restrict__ prob) {
  __shared__ float sp[kParams];
  __shared__ float snorm[2 * kIn];
  load_params(params, norm, sp, snorm);
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    float x[kIn];
    normalise(feats, s, snorm, x);
    float z = sp[kOffB2];
    for (int j = 0; j < kHid; ++j) {
      float a = sp[kOffB1 + j];
#pragma unroll
      for (int i = 0; i < kIn; ++i) a = fmaf(sp[j * kIn + i], x[i], a);
      z = fmaf(sp[kOffW2 + j], fmaxf(a, 0.f), z);
    }
    prob[s] = 1.f / (1.f + expf(-z));
  }
}

// Gradient of the mean weighted BCE over n samples.  Each wavefront takes 64
// samples at a time (one per lane), reduces every per-parameter contribution
// across its lanes in a fixed order, and keeps the running sums distributed
// over its lanes (parameter k lives in lane k % 64, slot k / 64).  The
// per-wavefront partials are summed in wavefront order by k_mlp_reduce.
__global__ void __launch_bounds__(kBlock)
k_mlp_grad(const float *__restrict__ params, const float *__restrict__ norm,
           const float *__restrict__ feats, const float *__restrict__ target, int64_t n,
           float pos_weight, float inv_n, float *__restrict__ partial,
           float *__restrict__ loss_partial) {
  __shared__ float sp[kParams];
  __shared__ float snorm[2 * kIn];
  load_params(params, norm, sp, snorm);
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t n_waves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  float acc[kSlots];
#pragma unroll
  for (int q = 0; q < kSlots; ++q) acc[q] = 0.f;
  float loss_acc = 0.f;
  for (int64_t base = wave * kWave; base < n; base += n_waves * kWave) {
    const int64_t s = base + lane;
    const bool valid = s < n;
    float x[kIn];
    if (valid) {
      normalise(feats, s, snorm, x);
    } else {
#pragma unroll
      for (int i = 0; i < kIn; ++i) x[i] = 0.f;
    }
    float h[kHid];
    float z = sp[kOffB2];
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
      float a = sp[kOffB1 + j];
#pragma unroll
      for (int i = 0; i < kIn; ++i) a = fmaf(sp[j * kIn + i], x[i], a);
      h[j] = fmaxf(a, 0.f);
      z = fmaf(sp[kOffW2 + j], h[j], z);
    }
    const float y = 1.f / (1.f + expf(-z));
    const float t = valid ? target[s] : 0.f;
    float g = 0.f;
    if (valid) {
      g = (pos_weight * t * (y - 1.f) + (1.f - t) * y) * inv_n;
      loss_acc += (pos_weight * t * softplus_f(-z) + (1.f - t) * softplus_f(z)) * inv_n;
    }
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
      const float gh = (h[j] > 0.f) ? g * sp[kOffW2 + j] : 0.f;
#pragma unroll
      for (int i = 0; i < kIn; ++i) {
        const int k = j * kIn + i;
        const float r = wave_sum(gh * x[i]);
        if (lane == (k & (kWave - 1))) acc[k / kWave] += r;
      }
      {
        const int k = kOffB1 + j;
        const float r = wave_sum(gh);
        if (lane == (k & (kWave - 1))) acc[k / kWave] += r;
      }
      {
        const int k = kOffW2 + j;
        const float r = wave_sum(g * h[j]);
        if (lane == (k & (kWave - 1))) acc[k / kWave] += r;
      }
    }
    {
      const float r = wave_sum(g);
      if (lane == (kOffB2 & (kWave - 1))) acc[kOffB2 / kWave] += r;
    }
  }
#pragma unroll
  for (int q = 0; q < kSlots; ++q) {
    const int k = q * kWave + lane;
    if (k < kParams) partial[wave * kParams + k] = acc[q];
  }
  loss_acc = wave_sum(loss_acc);
  if (lane == 0) loss_partial[wave] = loss_acc;
}

__global__ void k_mlp_reduce(const float *__restrict__ partial,
                             const float *__restrict__ loss_partial, int64_t n_waves,
                             float *__restrict__ grad, float *__restrict__ loss) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kParams) {
    float s = 0.f;
    for (int64_t w = 0; w < n_waves; ++w) s += partial[w * kParams + k];
    grad[k] = s;
  } else if (k == kParams && loss) {
    float s = 0.f;
    for (int64_t w = 0; w < n_waves; ++w) s += loss_partial[w];
    *loss = s;
  }
}

__global__ void k_adam(float *__restrict__ params, const float *__restrict__ grad,
                       float *__restrict__ m, float *__restrict__ v, int n, float lr, float b1,
                       float b2, float eps, float corr1, float corr2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float g = grad[k];
  const float mk = b1 * m[k] + (1.f - b1) * g;
  const float vk = b2 * v[k] + (1.f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  params[k] -= lr * (mk / corr1) / (sqrtf(vk / corr2) + eps);
}

// BF16 inference: weights, inputs and hidden activations rounded to bf16,
// products accumulated in FP32.  route[cand[i]] = prob >= threshold (or
// route[i] when cand is NULL).
__global__ void __launch_bounds__(kBlock)
k_mlp_route_bf16(const float *__restrict__ params, const float *__restrict__ norm,
                 const float *__restrict__ feats, int64_t n, float threshold,
                 const int32_t *__restrict__ cand, uint8_t *__restrict__ route,
                 float *__restrict__ prob) {
  __shared__ float sp[kParams];
  __shared__ float snorm[2 * kIn];
  load_params_bf16(params, norm, sp, snorm);
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    float f[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) f[i] = feats[s * kIn + i];
    const float y = mlp_prob_bf16(sp, snorm, f);
    if (prob) prob[s] = y;
    if (route) route[cand ? (int64_t)cand[s] : s] = (y >= threshold) ? 1 : 0;
  }
}

}  // namespace tc

extern "C" int tc_router_n_params(void) { return tc::kParams; }

extern "C" int tc_mlp_forward(const float *params, const float *norm, const float *feats,
                              int64_t n, float *prob, void *stream) {
  using namespace tc;
  TC_CHECK(params && norm && n >= 0 && (n == 0 || (feats && prob)), TC_ERR_INVALID,
           "invalid arguments to tc_mlp_forward");
  TC_TRY(ensure_init());
  if (n == 0) return TC_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0,
                     (hipStream_t)stream, params, norm, feats, n, prob);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_mlp_grad(const float *params, const float *norm, const float *feats,
                           const float *target, int64_t n, float pos_weight, float *grad,
                           float *loss, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(params && norm && grad && n > 0 && feats && target, TC_ERR_INVALID,
           "invalid arguments to tc_mlp_grad");
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  // a fixed number of wavefronts, so the reduction order does not depend on n
  const unsigned blocks = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)num_cus() * 2));
  const int64_t n_waves = (int64_t)blocks * (kBlock / kWave);
  Scratch partial, lossp;
  TC_TRY(partial.alloc(sizeof(float) * n_waves * kParams, s));
  TC_TRY(lossp.alloc(sizeof(float) * n_waves, s));
  hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kBlock), 0, s, params, norm, feats, target, n,
                     pos_weight, 1.0f / (float)n, partial.as<float>(), lossp.as<float>());
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     (const float *)partial.as<float>(), (const float *)lossp.as<float>(), n_waves,
                     grad, loss);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_mlp_adam(float *params, const float *grad, float *m, float *v, int n_params,
                           int step, float lr, float beta1, float beta2, float eps, void *stream) {
  using namespace tc;
  TC_CHECK(params && grad && m && v && n_params > 0 && step >= 1, TC_ERR_INVALID,
           "invalid arguments to tc_mlp_adam");
  TC_TRY(ensure_init());
  const float c1 = 1.f - powf(beta1, (float)step);
  const float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n_params + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     (hipStream_t)stream, params, grad, m, v, n_params, lr, beta1, beta2, eps, c1,
                     c2);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

extern "C" int tc_mlp_route_bf16(const float *params, const float *norm, const float *feats,
                                 int64_t n, float threshold, const int32_t *cand, uint8_t *route,
                                 float *prob, void *stream) {
  using namespace tc;
  TC_CHECK(params && norm && n >= 0 && (n == 0 || (feats && (route || prob))), TC_ERR_INVALID,
           "invalid arguments to tc_mlp_route_bf16");
  TC_TRY(ensure_init());
  if (n == 0) return TC_OK;
  hipLaunchKernelGGL(k_mlp_route_bf16, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0,
                     (hipStream_t)stream, params, norm, feats, n, threshold, cand, route, prob);
  TC_HIP(hipGetLastError());
  return TC_OK;
}
