// Stable LSD radix sort of 64-bit words (8-bit digits), reduce-then-scan form:
// per pass one histogram kernel, one exclusive scan of the digit-major
// (digit, tile) count table, and one scatter kernel.  A tile is 8192 keys
// for a 512-thread workgroup: its eight wavefronts own consecutive eighths,
// and a wavefront ranks its keys round by round with ballot "match" masks, so
// the order of equal digits is kept.  With uniform digits a tile writes runs
// of 32 keys (256 B) per digit.
#include <algorithm>

#include "tc_common.hpp"

namespace tc {

constexpr int kRadixBits = 8;
constexpr int kRadixDigits = 1 << kRadixBits;
constexpr int kRadixRounds = 16;                              // keys per lane
constexpr int kRadixBlock = 512;                              // threads per tile
constexpr int kRadixWaves = kRadixBlock / kWave;              // 8
constexpr int kRadixWaveKeys = kWave * kRadixRounds;          // 1024
constexpr int kRadixTile = kRadixWaveKeys * kRadixWaves;      // 8192
static_assert(kRadixBlock >= kRadixDigits, "the scatter assigns one digit per thread");

// Exclusive scan of one int per thread over the workgroup (*total = sum).
__device__ __forceinline__ int block_exclusive_scan_int(int x, int *wave_tot, int *total) {
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int inc = wave_inclusive_scan(x);
  if (lane == kWave - 1) wave_tot[w] = inc;
  __syncthreads();
  int before = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < kRadixWaves; ++i) {
    const int t = wave_tot[i];
    before += (i < w) ? t : 0;
    sum += t;
  }
  *total = sum;
  return before + inc - x;
}

__global__ void __launch_bounds__(kRadixBlock)
k_radix_hist(const uint64_t *__restrict__ keys, int64_t n_max,
             const unsigned long long *__restrict__ n_dev, int drop, uint64_t drop_key, int shift,
             int dmask, int64_t ntiles, int32_t *__restrict__ counts) {
  const int64_t n = n_dev ? min(n_max, (int64_t)*n_dev) : n_max;
  __shared__ int hist[kRadixDigits];
  for (int i = threadIdx.x; i < kRadixDigits; i += kRadixBlock) hist[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kRadixWaveKeys + lane;
#pragma unroll 4
  for (int r = 0; r < kRadixRounds; ++r) {
    const int64_t i = base + (int64_t)r * kWave;
    if (i < n) {
      const uint64_t key = keys[i];
      if (!(drop && key == drop_key)) atomicAdd(&hist[(int)(key >> shift) & dmask], 1);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadixDigits; d += kRadixBlock)
    counts[(int64_t)d * ntiles + blockIdx.x] = hist[d];
}

// Scatter of one tile: keys are ranked (stable) inside the tile, staged in
// LDS in digit order, then written out digit run by digit run, so that
// consecutive threads write consecutive addresses of the output.
__global__ void __launch_bounds__(kRadixBlock)
k_radix_scatter(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t n_max,
                const unsigned long long *__restrict__ n_dev, int drop, uint64_t drop_key,
                int shift, int dmask, int64_t ntiles, const int64_t *__restrict__ offs) {
  const int64_t n = n_dev ? min(n_max, (int64_t)*n_dev) : n_max;
  if ((int64_t)blockIdx.x * kRadixTile >= n) return;  // tile past the device count
  __shared__ int wpos[kRadixWaves][kRadixDigits];  // counts, then running positions
  __shared__ int tstart[kRadixDigits];                // first tile slot of each digit
  __shared__ int64_t gstart[kRadixDigits];            // first output slot of each digit
  __shared__ uint64_t stage[kRadixTile];
  __shared__ int wave_tot[kRadixWaves];
  __shared__ int s_kept;                              // keys of the tile that are sorted
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i = threadIdx.x; i < kRadixWaves * kRadixDigits; i += kRadixBlock)
    (&wpos[0][0])[i] = 0;
  const int64_t t0 = (int64_t)blockIdx.x * kRadixTile;
  const int64_t base = t0 + (int64_t)w * kRadixWaveKeys + lane;
  uint64_t k[kRadixRounds];
#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const int64_t i = base + (int64_t)r * kWave;
    k[r] = (i < n) ? in[i] : 0;
  }
  __syncthreads();

  // per-wavefront digit counts; peers = lanes of this round with the same digit
  uint64_t peer[kRadixRounds];
#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const bool valid = base + (int64_t)r * kWave < n && !(drop && k[r] == drop_key);
    const int d = (int)(k[r] >> shift) & dmask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < kRadixBits; ++bit) {
      const uint64_t m = __ballot((d >> bit) & 1);
      peers &= ((d >> bit) & 1) ? m : ~m;
    }
    peer[r] = valid ? peers : 0;
    if (valid && (peers & lt) == 0) atomicAdd(&wpos[w][d], __popcll(peers));
  }
  __syncthreads();

  // thread d < 256: the tile's digit-d slots start after all smaller digits;
  // the wavefronts' runs follow each other inside a digit (wavefront order is
  // input order)
  {
    const int d = threadIdx.x;
    const bool has_digit = d < kRadixDigits;
    int c[kRadixWaves], tot = 0;
#pragma unroll
    for (int ww = 0; ww < kRadixWaves; ++ww) {
      c[ww] = has_digit ? wpos[ww][d] : 0;
      tot += c[ww];
    }
    int kept;
    const int st = block_exclusive_scan_int(tot, wave_tot, &kept);
    if (threadIdx.x == 0) s_kept = kept;
    if (has_digit) {
      tstart[d] = st;
      gstart[d] = offs[(int64_t)d * ntiles + blockIdx.x];
      int p = st;
#pragma unroll
      for (int ww = 0; ww < kRadixWaves; ++ww) {
        wpos[ww][d] = p;
        p += c[ww];
      }
    }
  }
  __syncthreads();

#pragma unroll
  for (int r = 0; r < kRadixRounds; ++r) {
    const uint64_t peers = peer[r];
    const int d = (int)(k[r] >> shift) & dmask;
    if (peers) stage[wpos[w][d] + __popcll(peers & lt)] = k[r];
    wave_lds_sync();
    if (peers && (peers & lt) == 0) wpos[w][d] += __popcll(peers);
    wave_lds_sync();
  }
  __syncthreads();

  const int tile_kept = s_kept;
  for (int i = threadIdx.x; i < tile_kept; i += kRadixBlock) {
    const uint64_t key = stage[i];
    const int d = (int)(key >> shift) & dmask;
    out[gstart[d] + (i - tstart[d])] = key;
  }
}

size_t radix_sort_scratch_bytes(int64_t n, int bit_lo, int bit_hi) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  const int passes = (bit_hi - bit_lo + kRadixBits - 1) / kRadixBits;
  return arena_round(sizeof(int32_t) * ntiles * kRadixDigits) +
         arena_round(sizeof(int64_t) * (ntiles * kRadixDigits + 1)) +
         (size_t)passes * scan_scratch_bytes(ntiles * kRadixDigits);
}

int radix_sort_u64(uint64_t *a, uint64_t *b, int64_t n, int bit_lo, int bit_hi,
                   uint64_t **result, hipStream_t s, const uint64_t *drop,
                   const unsigned long long *n_dev) {
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
    // the first pass reads all n words (dropping *drop); later ones the kept count
    const bool first = shift == bit_lo;
    const int drop_now = (first && drop) ? 1 : 0;
    const uint64_t drop_key = drop ? *drop : 0;
    const unsigned long long *nd = first ? nullptr : n_dev;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kRadixBlock), 0, s,
                       (const uint64_t *)src, n, nd, drop_now, drop_key, shift, dmask, ntiles,
                       counts.as<int32_t>());
    TC_HIP(hipGetLastError());
    TC_TRY(exclusive_scan<int32_t>(counts.as<int32_t>(), offs.as<int64_t>(),
                                   ntiles * kRadixDigits, nullptr, s));
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)ntiles), dim3(kRadixBlock), 0, s,
                       (const uint64_t *)src, dst, n, nd, drop_now, drop_key, shift, dmask,
                       ntiles, (const int64_t *)offs.as<int64_t>());
    TC_HIP(hipGetLastError());
    std::swap(src, dst);
  }
  *result = src;
  return TC_OK;
}

}  // namespace tc
