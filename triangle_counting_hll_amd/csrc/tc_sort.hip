// Stable LSD radix sort of 64-bit words (8-bit digits), reduce-then-scan form:
// per pass one histogram kernel, one exclusive scan of the digit-major
// (digit, tile) count table, and one scatter kernel.  Inside a tile the four
// wavefronts own consecutive quarters, and a wavefront ranks its keys round
// by round with ballot "match" masks, so the order of equal digits is kept.
#include <algorithm>

#include "tc_common.hpp"

namespace tc {

constexpr int kRadixBits = 8;
constexpr int kRadixDigits = 1 << kRadixBits;
constexpr int kRadixRounds = 16;                              // keys per lane
constexpr int kRadixWaveKeys = kWave * kRadixRounds;         // 1024
constexpr int kRadixTile = kRadixWaveKeys * (kBlock / kWave);  // 4096

__global__ void __launch_bounds__(kBlock)
k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift, int dmask, int64_t ntiles,
             int32_t *__restrict__ counts) {
  __shared__ int hist[kRadixDigits];
  for (int i = threadIdx.x; i < kRadixDigits; i += kBlock) hist[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kRadixWaveKeys + lane;
#pragma unroll 4
  for (int r = 0; r < kRadixRounds; ++r) {
    const int64_t i = base + (int64_t)r * kWave;
    if (i < n) atomicAdd(&hist[(int)(keys[i] >> shift) & dmask], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadixDigits; d += kBlock)
    counts[(int64_t)d * ntiles + blockIdx.x] = hist[d];
}

__global__ void __launch_bounds__(kBlock)
k_radix_scatter(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t n, int shift,
                int dmask, int64_t ntiles, const int64_t *__restrict__ offs) {
  __shared__ int wcnt[kBlock / kWave][kRadixDigits];
  __shared__ int64_t wpos[kBlock / kWave][kRadixDigits];
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i = threadIdx.x; i < (kBlock / kWave) * kRadixDigits; i += kBlock)
    (&wcnt[0][0])[i] = 0;
  const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kRadixWaveKeys + lane;
  uint64_t k[kRadixRounds];
#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const int64_t i = base + (int64_t)r * kWave;
    k[r] = (i < n) ? in[i] : 0;
  }
  __syncthreads();

  // per-wavefront digit counts
#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const bool valid = base + (int64_t)r * kWave < n;
    const int d = (int)(k[r] >> shift) & dmask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kRadixBits; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1);
      peers &= ((d >> bit) & 1) ? m : ~m;
    }
    if (valid && (peers & lt) == 0) atomicAdd(&wcnt[w][d], __popcll(peers));
  }
  __syncthreads();

  // start of each (wavefront, digit) run in the output
  for (int d = threadIdx.x; d < kRadixDigits; d += kBlock) {
    int64_t p = offs[(int64_t)d * ntiles + blockIdx.x];
#pragma unroll
    for (int ww = 0; ww < kBlock / kWave; ++ww) {
      wpos[ww][d] = p;
      p += wcnt[ww][d];
    }
  }
  __syncthreads();

#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const bool valid = base + (int64_t)r * kWave < n;
    const int d = (int)(k[r] >> shift) & dmask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kRadixBits; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1);
      peers &= ((d >> bit) & 1) ? m : ~m;
    }
    if (valid) {
      const int64_t pos = wpos[w][d] + __popcll(peers & lt);
      out[pos] = k[r];
    }
    wave_lds_sync();
    if (valid && (peers & lt) == 0) wpos[w][d] += __popcll(peers);
    wave_lds_sync();
  }
}

int radix_sort_u64(uint64_t *a, uint64_t *b, int64_t n, int bit_lo, int bit_hi,
                   uint64_t **result, hipStream_t s) {
  *result = a;
  if (n <= 1 || bit_hi <= bit_lo) return TC_OK;
  const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  if (ntiles > INT32_MAX) {
    set_error(TC_ERR_INVALID, "radix sort of %lld keys is too large", (long long)n);
    return TC_ERR_INVALID;
  }
  Scratch counts, offs;
  TC_TRY(counts.alloc(sizeof(int32_t) * ntiles * kRadixDigits, s));
  TC_TRY(offs.alloc(sizeof(int64_t) * (ntiles * kRadixDigits + 1), s));
  uint64_t *src = a, *dst = b;
  for (int shift = bit_lo; shift < bit_hi; shift += kRadixBits) {
    // digits never reach past bit_hi, so higher bits do not perturb the order
    const int dmask = (1 << std::min(kRadixBits, bit_hi - shift)) - 1;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kBlock), 0, s,
                       (const uint64_t *)src, n, shift, dmask, ntiles, counts.as<int32_t>());
    TC_HIP(hipGetLastError());
    TC_TRY(exclusive_scan<int32_t>(counts.as<int32_t>(), offs.as<int64_t>(),
                                   ntiles * kRadixDigits, nullptr, s));
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)ntiles), dim3(kBlock), 0, s,
                       (const uint64_t *)src, dst, n, shift, dmask, ntiles,
                       (const int64_t *)offs.as<int64_t>());
    TC_HIP(hipGetLastError());
    std::swap(src, dst);
  }
  *result = src;
  return TC_OK;
}

}  // namespace tc
