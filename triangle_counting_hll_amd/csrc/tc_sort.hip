// Stable LSD radix sort of 64-bit words (8-bit digits), one-sweep form: one
// histogram kernel counts the digits of every pass in a single read of the
// input, then each pass is one scatter kernel whose tiles find their output
// offsets by decoupled look-back over the digit counts earlier tiles
// publish.  A tile is 8192 keys for a 512-thread workgroup: its eight
// wavefronts own consecutive eighths, and a wavefront ranks its keys round by
// round with ballot "match" masks, so the order of equal digits is kept.
// With uniform digits a tile writes runs of 32 keys (256 B) per digit.
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

// Digit histograms of every pass at once, over the kept words of the input:
// LDS counters per workgroup (kHistCopies copies of each, chosen by lane, so
// that lanes with one digit -- frequent in the high digits -- rarely update
// one address), then one global atomic per nonzero counter.
constexpr int kMaxPasses = 8;
constexpr int kHistCopies = 4;

__global__ void __launch_bounds__(kRadixBlock)
k_onesweep_hist(const uint64_t *__restrict__ keys, int64_t n,
                const unsigned long long *__restrict__ n_dev, int drop, uint64_t drop_key,
                int bit_lo, int bit_hi, unsigned long long *__restrict__ ghist) {
  if (n_dev) n = min(n, (int64_t)*n_dev);
  __shared__ unsigned int hist[kMaxPasses * kRadixDigits * kHistCopies];
  const int passes = (bit_hi - bit_lo + kRadixBits - 1) / kRadixBits;
  for (int i = threadIdx.x; i < passes * kRadixDigits * kHistCopies; i += kRadixBlock)
    hist[i] = 0;
  __syncthreads();
  const int copy = lane_id() & (kHistCopies - 1);
  for (int64_t i = (int64_t)blockIdx.x * kRadixBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRadixBlock) {
    const uint64_t key = keys[i];
    if (drop && key == drop_key) continue;
    for (int p = 0; p < passes; ++p) {
      const int sh = bit_lo + p * kRadixBits;
      const int dm = (1 << min(kRadixBits, bit_hi - sh)) - 1;
      atomicAdd(&hist[(p * kRadixDigits + ((int)(key >> sh) & dm)) * kHistCopies + copy], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kRadixDigits; i += kRadixBlock) {
    unsigned int t = 0;
#pragma unroll
    for (int c = 0; c < kHistCopies; ++c) t += hist[i * kHistCopies + c];
    if (t) atomicAdd(&ghist[i], (unsigned long long)t);
  }
}

// gbase[p][d] = number of kept keys whose pass-p digit is below d.
__global__ void k_onesweep_bases(const unsigned long long *__restrict__ ghist, int passes,
                                 int64_t *__restrict__ gbase) {
  const int p = threadIdx.x;
  if (p >= passes) return;
  int64_t run = 0;
  for (int d = 0; d < kRadixDigits; ++d) {
    gbase[p * kRadixDigits + d] = run;
    run += (int64_t)ghist[p * kRadixDigits + d];
  }
}

// Decoupled look-back status of one (tile, digit): the flag in the top two
// bits, the count below.  An aggregate is the tile's own count of the digit;
// a prefix is the count over this tile and every earlier one.
constexpr unsigned long long kStatusAgg = 1ull << 62;
constexpr unsigned long long kStatusPrefix = 2ull << 62;
constexpr unsigned long long kStatusValue = (1ull << 62) - 1;
// A look-back wait gives up after this many polls (each with a short sleep),
// so every wait ends; a tile that waits on an earlier one waits only on a
// workgroup that already took its ticket and publishes its aggregate before
// any wait of its own, so the limit is never reached in practice.  A stall
// is counted in *stalled.
constexpr int kLookbackPolls = 1 << 24;

__device__ __forceinline__ unsigned long long status_load(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One pass of the sort: a workgroup takes the next tile by ticket, ranks its
// keys (stable) inside the tile, publishes its digit counts, finds each
// digit's count over all earlier tiles by looking back through their
// published counts, then writes the tile's keys staged in LDS digit run by
// digit run, so that consecutive threads write consecutive addresses.
__global__ void __launch_bounds__(kRadixBlock)
k_onesweep_scatter(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, int64_t n_max,
                   const unsigned long long *__restrict__ n_dev, int drop, uint64_t drop_key,
                   int shift, int dmask, const int64_t *__restrict__ gbase,
                   unsigned long long *__restrict__ status, unsigned long long *__restrict__ ticket,
                   unsigned long long *__restrict__ stalled, uint8_t *__restrict__ run_flags,
                   int run_lo, uint64_t run_mask, int64_t *__restrict__ run_starts) {
  const int64_t n = n_dev ? min(n_max, (int64_t)*n_dev) : n_max;
  __shared__ int wpos[kRadixWaves][kRadixDigits];  // counts, then running positions
  __shared__ int tstart[kRadixDigits];                // first tile slot of each digit
  __shared__ int64_t gstart[kRadixDigits];            // first output slot of each digit
  __shared__ uint64_t stage[kRadixTile];
  __shared__ int wave_tot[kRadixWaves];
  __shared__ int s_kept;                              // keys of the tile that are sorted
  __shared__ int64_t s_tile;
  if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ticket, 1ull);
  __syncthreads();
  const int64_t tile = s_tile;
  if (tile * kRadixTile >= n) return;  // past the keys: no later tile has keys either
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int i = threadIdx.x; i < kRadixWaves * kRadixDigits; i += kRadixBlock)
    (&wpos[0][0])[i] = 0;
  const int64_t t0 = tile * kRadixTile;
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

  // thread d < 256: the tile's digit-d count is published at once, then the
  // count over earlier tiles is gathered; inside the tile the digit's slots
  // start after all smaller digits and the wavefronts' runs follow each other
  {
    const int d = threadIdx.x;
    const bool has_digit = d < kRadixDigits;
    int c[kRadixWaves], tot = 0;
#pragma unroll
    for (int ww = 0; ww < kRadixWaves; ++ww) {
      c[ww] = has_digit ? wpos[ww][d] : 0;
      tot += c[ww];
    }
    unsigned long long *my = status + tile * kRadixDigits + d;
    if (has_digit)
      status_store(my, (tile == 0 ? kStatusPrefix : kStatusAgg) | (unsigned long long)tot);
    int kept;
    const int st = block_exclusive_scan_int(tot, wave_tot, &kept);
    if (threadIdx.x == 0) s_kept = kept;
    if (has_digit) {
      unsigned long long before = 0;
      if (tile > 0) {
        for (int64_t j = tile - 1; j >= 0; --j) {
          const unsigned long long *q = status + j * kRadixDigits + d;
          unsigned long long v = status_load(q);
          int polls = 0;
          while ((v >> 62) == 0) {
            if (++polls > kLookbackPolls) {
              atomicAdd(stalled, 1ull);
              v = kStatusPrefix;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            v = status_load(q);
          }
          before += v & kStatusValue;
          if ((v >> 62) == 2) break;
        }
        status_store(my, kStatusPrefix | (before + (unsigned long long)tot));
      }
      tstart[d] = st;
      gstart[d] = gbase[d] + (int64_t)before;
      // run flags: the first slot of each (tile, digit) run is settled by
      // k_run_fixup, which needs its position
      if (run_flags) run_starts[tile * kRadixDigits + d] = tot ? gstart[d] : -1;
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
    const int64_t pos = gstart[d] + (i - tstart[d]);
    out[pos] = key;
    if (run_flags) {
      // inside the tile's digit run the previous sorted word is the previous
      // staged one; the run's first slot is written by k_run_fixup
      if (i > tstart[d])
        run_flags[pos] = ((key >> run_lo) & run_mask) != ((stage[i - 1] >> run_lo) & run_mask);
    }
  }
}

// Run flag of the first slot of every (tile, digit) run of the last pass:
// compared with the word before it, which another tile wrote.
__global__ void k_run_fixup(const uint64_t *__restrict__ sorted, const int64_t *__restrict__ starts,
                            int64_t n_starts, int run_lo, uint64_t run_mask,
                            uint8_t *__restrict__ run_flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_starts;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pos = starts[i];
    if (pos < 0) continue;
    run_flags[pos] = (pos == 0 || ((sorted[pos] >> run_lo) & run_mask) !=
                                      ((sorted[pos - 1] >> run_lo) & run_mask)) ? 1 : 0;
  }
}

// run_flags[i] = 0 for i in [kept, n): the slots past the sorted words.
__global__ void k_run_flags_tail(const unsigned long long *__restrict__ kept_dev, int64_t kept_host,
                                 int64_t n, uint8_t *__restrict__ run_flags) {
  const int64_t kept = kept_dev ? min(n, (int64_t)*kept_dev) : kept_host;
  for (int64_t i = kept + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    run_flags[i] = 0;
}

// Run flags of words that were not sorted (at most one word, or no key bits).
__global__ void k_run_flags_plain(const uint64_t *__restrict__ s, int64_t n,
                                  const unsigned long long *__restrict__ kept_dev, int run_lo,
                                  uint64_t run_mask, uint8_t *__restrict__ run_flags) {
  const int64_t kept = kept_dev ? min(n, (int64_t)*kept_dev) : n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    run_flags[i] = (i < kept && (i == 0 || ((s[i] >> run_lo) & run_mask) !=
                                               ((s[i - 1] >> run_lo) & run_mask))) ? 1 : 0;
}

__global__ void k_set_u64(unsigned long long *__restrict__ p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

size_t radix_sort_scratch_bytes(int64_t n, int bit_lo, int bit_hi, bool run_flags) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  return arena_round(sizeof(unsigned long long) * (kMaxPasses * kRadixDigits + kMaxPasses + 1)) +
         arena_round(sizeof(int64_t) * kMaxPasses * kRadixDigits) +
         arena_round(sizeof(unsigned long long) * ntiles * kRadixDigits) +
         (run_flags ? arena_round(sizeof(int64_t) * ntiles * kRadixDigits) : 0);
}

int radix_sort_u64(uint64_t *a, uint64_t *b, int64_t n, int bit_lo, int bit_hi,
                   uint64_t **result, hipStream_t s, const uint64_t *drop,
                   const unsigned long long *n_dev, uint8_t *run_flags) {
  *result = a;
  // runs compare every bit from bit_lo up, also those above bit_hi
  const int run_lo = bit_lo;
  const uint64_t run_mask = ~0ull >> bit_lo;
  if (n <= 1 || bit_hi <= bit_lo) {
    if (run_flags && n > 0) {
      hipLaunchKernelGGL(k_run_flags_plain, dim3(grid_for(n, kBlock, 4096)), dim3(kBlock), 0, s,
                         (const uint64_t *)a, n, n_dev, run_lo, run_mask, run_flags);
      TC_HIP(hipGetLastError());
    }
    return TC_OK;
  }
  const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
  const int passes = (bit_hi - bit_lo + kRadixBits - 1) / kRadixBits;
  if (ntiles > INT32_MAX || passes > kMaxPasses) {
    set_error(TC_ERR_INVALID, "radix sort of %lld keys over %d bits is out of range",
              (long long)n, bit_hi - bit_lo);
    return TC_ERR_INVALID;
  }
  // counters: digit histograms [passes][256], tile tickets [passes], stalls
  Scratch counters, bases, status, starts;
  const int64_t n_ctr = (int64_t)kMaxPasses * kRadixDigits + kMaxPasses + 1;
  TC_TRY(counters.alloc(sizeof(unsigned long long) * n_ctr, s));
  TC_TRY(bases.alloc(sizeof(int64_t) * kMaxPasses * kRadixDigits, s));
  TC_TRY(status.alloc(sizeof(unsigned long long) * ntiles * kRadixDigits, s));
  if (run_flags) TC_TRY(starts.alloc(sizeof(int64_t) * ntiles * kRadixDigits, s));
  unsigned long long *ghist = counters.as<unsigned long long>();
  unsigned long long *tickets = ghist + (int64_t)kMaxPasses * kRadixDigits;
  unsigned long long *stalled = tickets + kMaxPasses;
  // zero fills by kernel (graph-capture safe)
  hipLaunchKernelGGL(k_set_u64, dim3(grid_for(n_ctr, kBlock, 64)), dim3(kBlock), 0, s, ghist,
                     n_ctr, 0ull);
  const uint64_t drop_key = drop ? *drop : 0;
  hipLaunchKernelGGL(k_onesweep_hist, dim3((unsigned)std::min<int64_t>(ntiles, num_cus() * 4)),
                     dim3(kRadixBlock), 0, s, (const uint64_t *)a, n, drop ? nullptr : n_dev,
                     drop ? 1 : 0, drop_key, bit_lo, bit_hi, ghist);
  hipLaunchKernelGGL(k_onesweep_bases, dim3(1), dim3(kWave), 0, s,
                     (const unsigned long long *)ghist, passes, bases.as<int64_t>());
  TC_HIP(hipGetLastError());
  uint64_t *src = a, *dst = b;
  for (int p = 0; p < passes; ++p) {
    const int shift = bit_lo + p * kRadixBits;
    // digits never reach past bit_hi, so higher bits do not perturb the order
    const int dmask = (1 << std::min(kRadixBits, bit_hi - shift)) - 1;
    // with drop, the first pass reads all n words (dropping *drop) and later
    // ones the kept count; without, every pass reads *n_dev words
    const bool first = p == 0;
    const int drop_now = (first && drop) ? 1 : 0;
    const unsigned long long *nd = (first && drop) ? nullptr : n_dev;
    const bool last = p == passes - 1;
    hipLaunchKernelGGL(k_set_u64, dim3(grid_for(ntiles * kRadixDigits, kBlock, 4096)),
                       dim3(kBlock), 0, s, status.as<unsigned long long>(),
                       ntiles * kRadixDigits, 0ull);
    if (last && run_flags)  // tiles past the keys leave no run start (-1)
      hipLaunchKernelGGL(k_set_u64, dim3(grid_for(ntiles * kRadixDigits, kBlock, 4096)),
                         dim3(kBlock), 0, s, (unsigned long long *)starts.as<int64_t>(),
                         ntiles * kRadixDigits, ~0ull);
    hipLaunchKernelGGL(k_onesweep_scatter, dim3((unsigned)ntiles), dim3(kRadixBlock), 0, s,
                       (const uint64_t *)src, dst, n, nd, drop_now, drop_key, shift, dmask,
                       (const int64_t *)(bases.as<int64_t>() + (int64_t)p * kRadixDigits),
                       status.as<unsigned long long>(), tickets + p, stalled,
                       last ? run_flags : nullptr, run_lo, run_mask,
                       last ? starts.as<int64_t>() : nullptr);
    TC_HIP(hipGetLastError());
    std::swap(src, dst);
  }
  *result = src;
  if (run_flags) {
    hipLaunchKernelGGL(k_run_fixup, dim3(grid_for(ntiles * kRadixDigits, kBlock, 4096)),
                       dim3(kBlock), 0, s, (const uint64_t *)src,
                       (const int64_t *)starts.as<int64_t>(), ntiles * kRadixDigits, run_lo,
                       run_mask, run_flags);
    hipLaunchKernelGGL(k_run_flags_tail, dim3(grid_for(n, kBlock, 1024)), dim3(kBlock), 0, s,
                       n_dev, n, n, run_flags);
    TC_HIP(hipGetLastError());
  }
  return TC_OK;
}

}  // namespace tc
