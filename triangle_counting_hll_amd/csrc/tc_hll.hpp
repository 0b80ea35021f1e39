// HLL estimator and batched per-edge register statistics, shared by the
// sketch kernels and the exact pipeline's edge pass (the learned router's
// sampled-union feature).
#pragma once

#include "tc_common.hpp"

namespace tc {

__device__ __forceinline__ double hll_alpha(int m) {
  if (m == 16) return 0.673;
  if (m == 32) return 0.697;
  if (m == 64) return 0.709;
  return 0.7213 / (1.0 + 1.079 / (double)m);
}

// Estimate from the merged register statistics (src/hll.py:37-54):
// Z = sum_j 2^-r_j (zero registers included), zeros = #{j : r_j = 0}.
// TC_HLL_CPU_EXP (cpu_exp/triangle_hll_experiment.py:356-365): LinearCounting
// or the 2^32 large-range correction of the raw estimate, each returned as is.
// TC_HLL_PROTO32 (experiment1_gpu/gpu_experiment_fixed.py:236-249): LinearCounting
// only.
__device__ __forceinline__ double hll_estimate(int m, uint64_t zeros, double Z,
                                              int mode = TC_HLL_GPU) {
  const double md = (double)m;
  double E = hll_alpha(m) * md * md / Z;
  if (mode == TC_HLL_PROTO32) {
    if (E <= 2.5 * md && zeros > 0) E = md * log(md / (double)zeros);
    return E;
  }
  if (mode == TC_HLL_CPU_EXP) {
    if (E <= 2.5 * md && zeros > 0) return md * log(md / (double)zeros);
    const double two32 = 4294967296.0;
    if (E > two32 / 30.0) return -two32 * log(1.0 - E / two32);
    return E;
  }
  if (E <= 2.5 * md && zeros > 0) E = md * log(md / (double)zeros);
  const double two64 = 18446744073709551616.0;
  if (E > two64 / 30.0) E = -two64 * log(1.0 - E / two64);
  return E;
}

// 2^-r as a double built from its exponent bits (r <= 61).
__device__ __forceinline__ double pow2_neg(uint32_t r) {
  return __hiloint2double((int)((1023u - r) << 20), 0);
}

// ---------------------------------------------------------------------------
// Batched per-edge register statistics.
//
// A wavefront takes 64 edges at a time ("batch slots"), one slot per lane.  An
// edge's two register rows are read by a group of G = min(64, m / 16) lanes,
// lane s of the group reading the 16-byte chunks s, s + G, s + 2G, ... of both
// rows, so a wave step covers 64 / G edges and small sketches (m < 1024) keep
// every lane loading.  The group's sums are reduced with log2(G) butterfly
// steps and handed to the lane that owns the slot; after G steps every lane
// holds the statistics of its own edge and the 64 estimates (FP64 divide and
// log) run side by side instead of on one lane per edge.  Each lane sums its
// chunks in a fixed order and the butterfly is fixed, so results are
// deterministic.
// ---------------------------------------------------------------------------
struct RegStats {
  uint32_t zeros[3];  // [row a, row b, merged]; a and b only when kTriple
  double z[3];
};

// Number of zero bytes in a 32-bit word: a byte's top bit in t is set iff the
// byte is non-zero (adding 0x7F carries into bit 7 for any non-zero low 7 bits).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t w) {
  const uint32_t t = ((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w;
  return 4u - (uint32_t)__popc(t & 0x80808080u);
}

// Adds this lane's chunks of two register rows: Z terms per byte position in
// four accumulators (independent FP64 add chains), combined in a fixed order.
template <int G, bool kTriple>
__device__ __forceinline__ void group_partials(const uint8_t *__restrict__ ra,
                                               const uint8_t *__restrict__ rb, int m, int sub,
                                               RegStats &st) {
  constexpr int kQ = kTriple ? 3 : 1;  // statistics kept: [merged] or [a, b, merged]
  double z4[kQ][4] = {};
  uint32_t zeros[kQ] = {};
#pragma unroll 2
  for (int j = sub * 16; j < m; j += 16 * G) {
    const uint4 a = *(const uint4 *)(ra + j);
    const uint4 b = *(const uint4 *)(rb + j);
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, bw[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      zeros[kQ - 1] += zero_bytes(aw[q] | bw[q]);  // max is 0 iff both are 0
      if (kTriple) {
        zeros[0] += zero_bytes(aw[q]);
        zeros[1] += zero_bytes(bw[q]);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t x = (aw[q] >> (8 * k)) & 0xffu, y = (bw[q] >> (8 * k)) & 0xffu;
        z4[kQ - 1][k] += pow2_neg(max(x, y));
        if (kTriple) {
          z4[0][k] += pow2_neg(x);
          z4[1][k] += pow2_neg(y);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int o = kTriple ? q : 2;
    st.zeros[o] += zeros[q];
    st.z[o] += (z4[q][0] + z4[q][1]) + (z4[q][2] + z4[q][3]);
  }
}

// Statistics of the edge in this lane's batch slot over the first `count`
// registers of its two rows (rows `stride` bytes apart).  row_a / row_b: this
// lane's slot's register rows (any value for unused slots).  n_valid: slots in
// use.
template <int G, bool kTriple>
__device__ __forceinline__ RegStats batch_stats(const uint8_t *__restrict__ regs, int stride,
                                                int count, int32_t row_a, int32_t row_b,
                                                int n_valid) {
  constexpr int kEpw = kWave / G;  // edges per wave step
  const int lane = lane_id();
  const int sub = lane % G, grp = lane / G;
  RegStats mine = {};
  for (int t = 0; t < G; ++t) {
    const int src = t * kEpw + grp;  // slot handled by this lane's group
    const int32_t a = __shfl(row_a, src, kWave), b = __shfl(row_b, src, kWave);
    RegStats st = {};
    if (src < n_valid)
      group_partials<G, kTriple>(regs + (int64_t)a * stride, regs + (int64_t)b * stride, count,
                                 sub, st);
#pragma unroll
    for (int q = kTriple ? 0 : 2; q < 3; ++q) {
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) {
        st.zeros[q] += __shfl_xor(st.zeros[q], off, kWave);
        st.z[q] += __shfl_xor(st.z[q], off, kWave);
      }
      // slot `lane` is handled at step lane / kEpw by group lane % kEpw
      const int from = (lane % kEpw) * G;
      const uint32_t zz = __shfl(st.zeros[q], from, kWave);
      const double zv = __shfl(st.z[q], from, kWave);
      if (lane / kEpw == t) {
        mine.zeros[q] = zz;
        mine.z[q] = zv;
      }
    }
  }
  return mine;
}

// Lanes per edge for m = 2^p registers: 16 bytes per lane per step, at most a wave.
inline int group_lanes(int p) { return p >= 10 ? 64 : 1 << (p - 4); }

#define TC_DISPATCH_GROUP(p, G, ...)                                 \
  switch (group_lanes(p)) {                                          \
    case 1: { constexpr int G = 1; __VA_ARGS__; } break;             \
    case 2: { constexpr int G = 2; __VA_ARGS__; } break;             \
    case 4: { constexpr int G = 4; __VA_ARGS__; } break;             \
    case 8: { constexpr int G = 8; __VA_ARGS__; } break;             \
    case 16: { constexpr int G = 16; __VA_ARGS__; } break;           \
    case 32: { constexpr int G = 32; __VA_ARGS__; } break;           \
    default: { constexpr int G = 64; __VA_ARGS__; } break;           \
  }

}  // namespace tc
