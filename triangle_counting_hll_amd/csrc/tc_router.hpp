// The learned router's network and features as device functions, shared by
// the stand-alone router kernels (features, BF16 inference) and the exact
// pipeline's edge pass, which evaluates the network per candidate edge.
//
// Network: 6 -> 64 -> 1, ReLU, sigmoid.  Parameter vector (float32, kParams):
//   W1[j][i] at j*6 + i (j < 64, i < 6), b1[j] at 384 + j, w2[j] at 448 + j,
//   b2 at 512.
// Normalisation vector (float32, 12 entries): mean[6] then inv_std[6]; the
// network sees x = (feature - mean) * inv_std.
#pragma once

#include "tc_common.hpp"

namespace tc {

constexpr int kIn = 6;
constexpr int kHid = 64;
constexpr int kOffB1 = kHid * kIn;    // 384
constexpr int kOffW2 = kOffB1 + kHid;  // 448
constexpr int kOffB2 = kOffW2 + kHid;  // 512
constexpr int kParams = kOffB2 + 1;    // 513

// float32 -> bfloat16, round to nearest even (NaN not expected)
__device__ __forceinline__ float round_bf16(float f) {
  uint32_t b = __float_as_uint(f);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return __uint_as_float(b & 0xFFFF0000u);
}

// The six features of an edge (u, v):
//   f0 = log2(1 + deg u), f1 = log2(1 + deg v), f2 = f0 - f1,
//   f3 = log2(1 + card u), f4 = log2(1 + card v)  (single-sketch estimates),
//   f5 = log2(1 + sampled union estimate).
__device__ __forceinline__ void router_features(int64_t deg_u, int64_t deg_v, float card_u,
                                                float card_v, double sampled_union,
                                                float f[kIn]) {
  const float f0 = log2f(1.0f + (float)deg_u);
  const float f1 = log2f(1.0f + (float)deg_v);
  f[0] = f0;
  f[1] = f1;
  f[2] = f0 - f1;
  f[3] = log2f(1.0f + card_u);
  f[4] = log2f(1.0f + card_v);
  f[5] = (float)log2(1.0 + sampled_union);
}

// Loads the parameters for BF16 inference into LDS (weights rounded to bf16,
// biases kept in FP32) with every thread of the workgroup; ends with a
// workgroup barrier.
__device__ __forceinline__ void load_params_bf16(const float *__restrict__ params,
                                                 const float *__restrict__ norm, float *sp,
                                                 float *snorm) {
  for (int k = threadIdx.x; k < kParams; k += blockDim.x) {
    const float w = params[k];
    const bool weight = k < kOffB1 || (k >= kOffW2 && k < kOffB2);
    sp[k] = weight ? round_bf16(w) : w;
  }
  for (int k = threadIdx.x; k < 2 * kIn; k += blockDim.x) snorm[k] = norm[k];
  __syncthreads();
}

// BF16 inference of one sample from LDS parameters (load_params_bf16):
// normalised inputs and hidden activations rounded to bf16, products
// accumulated in FP32.  Returns the sigmoid output.
__device__ __forceinline__ float mlp_prob_bf16(const float *sp, const float *snorm,
                                               const float f[kIn]) {
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) x[i] = round_bf16((f[i] - snorm[i]) * snorm[kIn + i]);
  float z = 0.f;
  for (int j = 0; j < kHid; ++j) {
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < kIn; ++i) a = fmaf(sp[j * kIn + i], x[i], a);
    a += sp[kOffB1 + j];
    z = fmaf(sp[kOffW2 + j], round_bf16(fmaxf(a, 0.f)), z);
  }
  z += sp[kOffB2];
  return 1.f / (1.f + expf(-z));
}

}  // namespace tc
