// (a) Device CSR build with (degree, id) orientation, and the degree-bucket sort.
//
// Pipeline (all on the device):
//   canonical (min,max) keys packed in 2 x bit_length(max id) bits -> LSD
//   radix sort -> adjacent-unique
//   -> presence flags + scan = dense relabel by sorted unique endpoint
//   -> undirected degrees (first ids per run of equal ids; second ids grouped
//   by a radix sort on their high bits and counted in LDS histograms)
//   -> orientation
//   -> oriented (src, dst) keys radix-sorted: the rows in order, each sorted
//   -> row_ptr by a lower bound per vertex over the sorted keys.
// The output equals the lexsort-based loader of experiment/exp_gpu.py:301-365.
#include <algorithm>
#include <climits>
#include <vector>

#include "tc_common.hpp"

namespace tc {

constexpr uint64_t kEmptyKey = ~0ull;

// Largest endpoint of the canonical non-loop pairs; flags negative ids.
__global__ void k_canon_max(const int32_t *__restrict__ pairs, int64_t m, int *__restrict__ max_id,
                            int *__restrict__ bad) {
  int local_max = -1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = pairs[2 * i], y = pairs[2 * i + 1];
    if (x < 0 || y < 0) {
      atomicOr(bad, 1);
      continue;
    }
    if (x != y) local_max = max(local_max, max(x, y));
  }
  for (int off = 32; off > 0; off >>= 1) local_max = max(local_max, __shfl_xor(local_max, off));
  // one same-address atomic per workgroup: thousands per wavefront serialise
  __shared__ int wave_max[kBlock / kWave];
  if (lane_id() == 0) wave_max[threadIdx.x / kWave] = local_max;
  __syncthreads();
  if (threadIdx.x == 0) {
    int bm = -1;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) bm = max(bm, wave_max[w]);
    if (bm >= 0) atomicMax(max_id, bm);
  }
}

// (min << B | max) for a non-loop pair, all-ones (sorts last) for a loop.
__global__ void k_canon_pack(const int32_t *__restrict__ pairs, int64_t m, int B,
                             uint64_t *__restrict__ key) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t x = pairs[2 * i], y = pairs[2 * i + 1];
    key[i] = (x == y) ? kEmptyKey
                      : (((uint64_t)(uint32_t)min(x, y) << B) | (uint64_t)(uint32_t)max(x, y));
  }
}

// First of each run of equal sorted keys (loops excluded).
__global__ void k_unique_flags(const uint64_t *__restrict__ k, int64_t m, int B,
                               uint8_t *__restrict__ flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (k[i] != kEmptyKey && (i == 0 || k[i] != k[i - 1])) ? 1 : 0;
}

// Unique canonical pairs, ascending by (min, max), as (min << 32 | max).
__global__ void k_unpack_unique(const uint64_t *__restrict__ k, const int32_t *__restrict__ pos,
                                const int64_t *__restrict__ n_p, int B,
                                uint64_t *__restrict__ out) {
  const int64_t n = *n_p;
  const uint64_t lo_mask = (1ull << B) - 1;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t w = k[pos[j]];
    out[j] = ((w >> B) << 32) | (w & lo_mask);
  }
}

__global__ void k_mark_present(const uint64_t *__restrict__ uniq, int64_t nu,
                               uint8_t *__restrict__ present) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nu;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = uniq[i];
    present[k >> 32] = 1;
    present[k & 0xffffffffu] = 1;
  }
}

// Relabel to dense ids in place (key -> dense a<<32 | dense b) and count the
// first ids' share of the undirected degrees.  The keys are sorted, so equal
// first ids sit in consecutive lanes: one atomic per run of them (hub rows
// would otherwise serialise thousands of same-address atomics).  The second
// ids are spread out; they are written to `second` and counted by bucket
// (k_bucket_degrees) instead of by one device-scope atomic each.
__global__ void k_relabel_degree(uint64_t *__restrict__ uniq, int64_t nu,
                                 const int64_t *__restrict__ nidx, int *__restrict__ udeg,
                                 uint64_t *__restrict__ second) {
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // the loop bound is uniform per wavefront, so every lane joins the ballot
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1));
       base < nu; base += stride) {
    const int64_t i = base + lane;
    const bool ok = i < nu;
    uint32_t a = 0xffffffffu, b = 0;
    if (ok) {
      const uint64_t k = uniq[i];
      a = (uint32_t)nidx[k >> 32];
      b = (uint32_t)nidx[k & 0xffffffffu];
      uniq[i] = ((uint64_t)a << 32) | b;
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)a, 1);
    const bool head = ok && (lane == 0 || prev != a);
    const uint64_t heads = __ballot(head);
    if (head) {
      const uint64_t above = heads & ~((2ull << lane) - 1);  // lane 63: no lanes above
      const int n_valid = (int)min<int64_t>(kWave, nu - base);
      const int next = above ? (__ffsll((unsigned long long)above) - 1) : n_valid;
      atomicAdd(&udeg[a], next - lane);
    }
    if (ok) second[i] = b;
  }
}

// The second ids' share of the degrees.  The ids are grouped by bucket (their
// bits from kDegBucketBits up, radix-sorted); bounds[j] is the first position
// of bucket j.  A workgroup takes a fixed chunk of the grouped ids (buckets
// are skewed: the first ids of a power-law graph are its hubs), counts each
// bucket segment of it in an LDS histogram of 2^kDegBucketBits counters and
// adds the nonzero counts to udeg: one atomic per (chunk, id) instead of one
// per edge, on consecutive addresses.
constexpr int kDegBucketBits = 14;  // 16384 counters, 64 KiB of LDS
constexpr int kDegBucketBlock = 1024;
constexpr int64_t kDegChunk = int64_t(1) << 18;

__global__ void k_bucket_bounds(const uint64_t *__restrict__ ids, int64_t nu, int64_t n_buckets,
                                int64_t *__restrict__ bounds) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= n_buckets;
       j += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = (uint64_t)j << kDegBucketBits;
    int64_t lo = 0, hi = nu;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (ids[mid] < key) lo = mid + 1; else hi = mid;
    }
    bounds[j] = lo;
  }
}

__global__ void __launch_bounds__(kDegBucketBlock)
k_bucket_degrees(const uint64_t *__restrict__ ids, int64_t nu, const int64_t *__restrict__ bounds,
                 int64_t n, int *__restrict__ udeg) {
  __shared__ int hist[1 << kDegBucketBits];
  const uint64_t low = (1ull << kDegBucketBits) - 1;
  const int64_t c1 = min(nu, ((int64_t)blockIdx.x + 1) * kDegChunk);
  int64_t pos = (int64_t)blockIdx.x * kDegChunk;
  while (pos < c1) {  // uniform: every thread reads the same ids[pos]
    const int64_t j = (int64_t)(ids[pos] >> kDegBucketBits);
    const int64_t seg_end = min(c1, bounds[j + 1]);
    for (int k = threadIdx.x; k < (1 << kDegBucketBits); k += kDegBucketBlock) hist[k] = 0;
    __syncthreads();
    for (int64_t i = pos + threadIdx.x; i < seg_end; i += kDegBucketBlock)
      atomicAdd(&hist[(int)(ids[i] & low)], 1);
    __syncthreads();
    const int64_t v0 = j << kDegBucketBits;
    for (int k = threadIdx.x; k < (1 << kDegBucketBits); k += kDegBucketBlock)
      if (hist[k] && v0 + k < n) atomicAdd(&udeg[v0 + k], hist[k]);
    __syncthreads();
    pos = seg_end;
  }
}

// u -> v iff (deg u, u) < (deg v, v); keys become (src << 32 | dst).  The
// forward degrees come later from the sorted keys (k_row_starts), so there is
// no per-edge atomic here.
__global__ void k_orient(uint64_t *__restrict__ uniq, int64_t nu, const int *__restrict__ udeg) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nu;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = uniq[i];
    uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)(k & 0xffffffffu);
    int da = udeg[a], db = udeg[b];
    bool a_first = (da < db) || (da == db && a < b);
    uint32_t s = a_first ? a : b, d = a_first ? b : a;
    uniq[i] = ((uint64_t)s << 32) | d;
  }
}

// row_ptr[v] = number of sorted (s << B | d) keys with s < v, for v in [0, n]
// (a lower bound per vertex; vertices without forward edges get the next
// row's start).
__global__ void k_row_starts(const uint64_t *__restrict__ keys, int64_t E, int B, int64_t n,
                             int64_t *__restrict__ row_ptr) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n;
       v += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = E;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((int64_t)(keys[mid] >> B) < v) lo = mid + 1; else hi = mid;
    }
    row_ptr[v] = lo;
  }
}

// Oriented keys (s << 32 | d) repacked as (s << B | d) for a radix sort on
// 2B bits (dense ids are below 2^B).
__global__ void k_repack_oriented(uint64_t *__restrict__ keys, int64_t E, int B) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    keys[i] = ((k >> 32) << B) | (k & 0xffffffffu);
  }
}

// Sorted (s << B | d) keys are the CSR in order: col = d, edge_u = s.
__global__ void k_unpack_rows(const uint64_t *__restrict__ keys, int64_t E, int B,
                              int32_t *__restrict__ col, int32_t *__restrict__ edge_u,
                              int32_t *__restrict__ edge_v) {
  const uint64_t mask = (1ull << B) - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    const int32_t d = (int32_t)(k & mask);
    col[i] = d;
    if (edge_u) edge_u[i] = (int32_t)(k >> B);
    if (edge_v) edge_v[i] = d;
  }
}

__global__ void k_finish_arrays(const int64_t *__restrict__ row_ptr64, int64_t n,
                                const int32_t *__restrict__ node_list,
                                const int *__restrict__ udeg,
                                void *__restrict__ row_ptr_out, int row_ptr_64,
                                int32_t *__restrict__ degrees, int64_t *__restrict__ node_ids,
                                int32_t *__restrict__ udeg_out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (row_ptr_out) {
      if (row_ptr_64)
        ((int64_t *)row_ptr_out)[i] = row_ptr64[i];
      else
        ((int32_t *)row_ptr_out)[i] = (int32_t)row_ptr64[i];
    }
    if (i < n) {
      if (degrees) degrees[i] = (int32_t)(row_ptr64[i + 1] - row_ptr64[i]);
      if (node_ids) node_ids[i] = node_list[i];
      if (udeg_out) udeg_out[i] = udeg[i];
    }
  }
}

}  // namespace tc

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
struct tc_csr_build {
  hipStream_t stream = nullptr;
  int64_t n = 0, E = 0;
  uint64_t *keys = nullptr;      // oriented (src<<32 | dst), dense ids
  int *udeg = nullptr;
  int64_t *row_ptr64 = nullptr;  // n+1, filled by tc_csr_build_finish
  int32_t *node_list = nullptr;  // n original ids
  ~tc_csr_build() {
    for (void *p : {(void *)keys, (void *)udeg, (void *)row_ptr64, (void *)node_list})
      if (p) (void)hipFree(p);
  }
};

extern "C" void tc_csr_build_free(tc_csr_build *b) { delete b; }

extern "C" int tc_csr_build_begin(const int32_t *pairs, int64_t m, void *stream,
                                  tc_csr_build **out, int64_t *n_nodes, int64_t *n_edges) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(out && n_nodes && n_edges, TC_ERR_INVALID, "null output pointer");
  TC_CHECK(m >= 0 && (m == 0 || pairs), TC_ERR_INVALID, "invalid pair array (m=%lld)",
           (long long)m);
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  tc_csr_build *b = new tc_csr_build();
  b->stream = s;
  *out = nullptr;
  auto fail = [&](int rc) {
    delete b;
    return rc;
  };

  // 1. largest id among canonical non-loop pairs (sizes the packed keys)
  Scratch counters;
  int rc = counters.alloc(64, s);
  if (rc) return fail(rc);
  if (hipMemsetAsync(counters.ptr, 0, 64, s) != hipSuccess) {
    set_error(TC_ERR_HIP, "memset failed");
    return fail(TC_ERR_HIP);
  }
  unsigned long long *n_uniq_d = (unsigned long long *)counters.ptr;
  int *max_id = (int *)((char *)counters.ptr + 8);
  int *bad = (int *)((char *)counters.ptr + 12);
  int minus1 = -1;
  if (hipMemcpyAsync(max_id, &minus1, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error(TC_ERR_HIP, "memcpy failed");
    return fail(TC_ERR_HIP);
  }
  if (m > 0)
    hipLaunchKernelGGL(k_canon_max, dim3(grid_for(m, kBlock, 2048)), dim3(kBlock), 0, s, pairs,
                       m, max_id, bad);
  struct {
    unsigned long long nu;
    int max_id;
    int bad;
  } h;
  if (hipMemcpyAsync(&h, counters.ptr, 16, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "id scan failed: %s", hipGetErrorString(hipGetLastError()));
    return fail(TC_ERR_HIP);
  }
  if (h.bad) {
    set_error(TC_ERR_INVALID, "edge list holds negative node ids");
    return fail(TC_ERR_INVALID);
  }
  // 2. canonical pairs packed as (min << B | max), B = bit length of the
  //    largest id, loops as all-ones; radix sort; keep the first of each run
  int B = 1;
  while (B < 32 && (int64_t(1) << B) <= (int64_t)h.max_id) ++B;
  Scratch packed, tmp, flags, pos;
  if ((rc = packed.alloc(sizeof(uint64_t) * std::max<int64_t>(m, 1), s)) ||
      (rc = tmp.alloc(sizeof(uint64_t) * std::max<int64_t>(m, 1), s)) ||
      (rc = flags.alloc(std::max<int64_t>(m, 1), s)) ||
      (rc = pos.alloc(sizeof(int32_t) * std::max<int64_t>(m, 1), s)))
    return fail(rc);
  if (hipMalloc(&b->keys, sizeof(uint64_t) * std::max<int64_t>(m, 1)) != hipSuccess) {
    set_error(TC_ERR_OOM, "cannot allocate %lld edge keys", (long long)m);
    return fail(TC_ERR_OOM);
  }
  int64_t nu = 0;
  if (m > 0 && h.max_id >= 0) {
    hipLaunchKernelGGL(k_canon_pack, dim3(grid_for(m, kBlock, 8192)), dim3(kBlock), 0, s, pairs,
                       m, B, packed.as<uint64_t>());
    uint64_t *sorted = nullptr;
    if ((rc = radix_sort_u64(packed.as<uint64_t>(), tmp.as<uint64_t>(), m, 0, 2 * B, &sorted, s)))
      return fail(rc);
    hipLaunchKernelGGL(k_unique_flags, dim3(grid_for(m, kBlock, 8192)), dim3(kBlock), 0, s,
                       (const uint64_t *)sorted, m, B, flags.as<uint8_t>());
    if ((rc = compact_indices(flags.as<uint8_t>(), m, pos.as<int32_t>(), (int64_t *)n_uniq_d, s)))
      return fail(rc);
    hipLaunchKernelGGL(k_unpack_unique, dim3(grid_for(m, kBlock, 8192)), dim3(kBlock), 0, s,
                       (const uint64_t *)sorted, (const int32_t *)pos.as<int32_t>(),
                       (const int64_t *)n_uniq_d, B, b->keys);
    int64_t nu_h = 0;
    if (hipMemcpyAsync(&nu_h, n_uniq_d, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error(TC_ERR_HIP, "dedup failed: %s", hipGetErrorString(hipGetLastError()));
      return fail(TC_ERR_HIP);
    }
    nu = nu_h;
  }
  const int64_t span = nu > 0 ? (int64_t)h.max_id + 1 : 0;  // 0 when no edges survive
  b->E = nu;

  Scratch present, nidx;
  rc = present.alloc(std::max<int64_t>(span, 1), s);
  if (rc) return fail(rc);
  rc = nidx.alloc(sizeof(int64_t) * (span + 1), s);
  if (rc) return fail(rc);
  if (span > 0) {
    if (hipMemsetAsync(present.ptr, 0, span, s) != hipSuccess) {
      set_error(TC_ERR_HIP, "memset failed");
      return fail(TC_ERR_HIP);
    }
    hipLaunchKernelGGL(k_mark_present, dim3(grid_for(nu, kBlock, 8192)), dim3(kBlock), 0, s,
                       (const uint64_t *)b->keys, nu, present.as<uint8_t>());
  }
  rc = exclusive_scan<uint8_t>(present.as<uint8_t>(), nidx.as<int64_t>(), span, nullptr, s);
  if (rc) return fail(rc);
  int64_t n = 0;
  if (hipMemcpyAsync(&n, nidx.as<int64_t>() + span, sizeof(int64_t), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "relabel readback failed");
    return fail(TC_ERR_HIP);
  }
  b->n = n;
  if (hipMalloc(&b->udeg, sizeof(int) * std::max<int64_t>(n, 1)) != hipSuccess ||
      hipMalloc(&b->row_ptr64, sizeof(int64_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&b->node_list, sizeof(int32_t) * std::max<int64_t>(n, 1)) != hipSuccess) {
    set_error(TC_ERR_OOM, "cannot allocate CSR build state for %lld nodes", (long long)n);
    return fail(TC_ERR_OOM);
  }
  if (hipMemsetAsync(b->udeg, 0, sizeof(int) * std::max<int64_t>(n, 1), s) != hipSuccess) {
    set_error(TC_ERR_HIP, "memset failed");
    return fail(TC_ERR_HIP);
  }
  if (span > 0) {
    rc = compact_indices(present.as<uint8_t>(), span, b->node_list, nullptr, s);
    if (rc) return fail(rc);
  }
  if (nu > 0) {
    // the canonical-sort buffers (m >= nu words each) hold the second ids
    uint64_t *second = packed.as<uint64_t>();
    hipLaunchKernelGGL(k_relabel_degree, dim3(grid_for(nu, kBlock, 8192)), dim3(kBlock), 0, s,
                       b->keys, nu, (const int64_t *)nidx.as<int64_t>(), b->udeg, second);
    int nb_bits = 1;
    while ((int64_t(1) << nb_bits) < n) ++nb_bits;
    uint64_t *grouped = second;
    if (nb_bits > kDegBucketBits &&
        (rc = radix_sort_u64(second, tmp.as<uint64_t>(), nu, kDegBucketBits, nb_bits, &grouped,
                             s)))
      return fail(rc);
    const int64_t n_buckets = (n + (int64_t(1) << kDegBucketBits) - 1) >> kDegBucketBits;
    Scratch bounds;
    if ((rc = bounds.alloc(sizeof(int64_t) * (n_buckets + 1), s))) return fail(rc);
    hipLaunchKernelGGL(k_bucket_bounds, dim3(grid_for(n_buckets + 1)), dim3(kBlock), 0, s,
                       (const uint64_t *)grouped, nu, n_buckets, bounds.as<int64_t>());
    hipLaunchKernelGGL(k_bucket_degrees, dim3((unsigned)((nu + kDegChunk - 1) / kDegChunk)),
                       dim3(kDegBucketBlock), 0, s, (const uint64_t *)grouped, nu,
                       (const int64_t *)bounds.as<int64_t>(), n, b->udeg);
    hipLaunchKernelGGL(k_orient, dim3(grid_for(nu, kBlock, 8192)), dim3(kBlock), 0, s, b->keys,
                       nu, (const int *)b->udeg);
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "CSR build phase 1 failed: %s", hipGetErrorString(hipGetLastError()));
    return fail(TC_ERR_HIP);
  }
  *n_nodes = n;
  *n_edges = nu;
  *out = b;
  return TC_OK;
}

extern "C" int tc_csr_build_finish(tc_csr_build *b, void *row_ptr, int row_ptr_64,
                                   int32_t *col_idx, int32_t *degrees, int64_t *node_ids,
                                   int32_t *edge_u, int32_t *edge_v, int32_t *undirected_deg,
                                   void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(b, TC_ERR_INVALID, "null CSR build state");
  TC_CHECK(b->E == 0 || col_idx, TC_ERR_INVALID, "col_idx is required");
  TC_CHECK(row_ptr_64 || b->E <= INT32_MAX, TC_ERR_INVALID,
           "%lld oriented edges do not fit int32 row offsets; use int64", (long long)b->E);
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = b->n, E = b->E;
  int B = 1;
  while ((int64_t(1) << B) < n) ++B;
  const uint64_t *sorted_keys = b->keys;  // read only when E > 0
  Scratch tmp;
  if (E > 0) {
    // rows in order and sorted inside: one radix sort of the oriented keys
    TC_TRY(tmp.alloc(sizeof(uint64_t) * E, s));
    hipLaunchKernelGGL(k_repack_oriented, dim3(grid_for(E, kBlock, 16384)), dim3(kBlock), 0, s,
                       b->keys, E, B);
    TC_HIP(hipGetLastError());
    uint64_t *sorted = nullptr;
    TC_TRY(radix_sort_u64(b->keys, tmp.as<uint64_t>(), E, 0, 2 * B, &sorted, s));
    hipLaunchKernelGGL(k_unpack_rows, dim3(grid_for(E, kBlock, 16384)), dim3(kBlock), 0, s,
                       (const uint64_t *)sorted, E, B, col_idx, edge_u, edge_v);
    TC_HIP(hipGetLastError());
    sorted_keys = sorted;
  }
  // row offsets (and forward degrees) from the sorted keys
  hipLaunchKernelGGL(k_row_starts, dim3(grid_for(n + 1)), dim3(kBlock), 0, s, sorted_keys, E, B,
                     n, b->row_ptr64);
  TC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finish_arrays, dim3(grid_for(n + 1)), dim3(kBlock), 0, s,
                     (const int64_t *)b->row_ptr64, n, (const int32_t *)b->node_list,
                     (const int *)b->udeg, row_ptr, row_ptr_64, degrees, node_ids,
                     undirected_deg);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

// ---------------------------------------------------------------------------
// Degree-bucket sort: bucket(u) = ceil(log2(deg+(u) + 1)), stable in u,
// on top of the generic stable bucket sort of tc_common.
// ---------------------------------------------------------------------------
namespace tc {
constexpr int kDegBuckets = 33;

template <class I>
__global__ void k_degree_keys(const I *__restrict__ rp, int64_t n, uint8_t *__restrict__ key) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t d = (uint64_t)((int64_t)rp[u + 1] - (int64_t)rp[u]);
    key[u] = (uint8_t)(d == 0 ? 0 : 64 - __clzll((long long)d));  // bit length of d
  }
}
}  // namespace tc

extern "C" int tc_degree_buckets(const void *row_ptr, int row_ptr_64, int64_t n_nodes,
                                 int64_t *bucket_start, int32_t *perm, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(row_ptr && bucket_start && perm && n_nodes >= 0, TC_ERR_INVALID,
           "invalid arguments to tc_degree_buckets");
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  Scratch keys;
  TC_TRY(keys.alloc(std::max<int64_t>(n_nodes, 1), s));
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    hipLaunchKernelGGL(k_degree_keys<I>, dim3(grid_for(n_nodes, kBlock, 16384)), dim3(kBlock), 0,
                       s, (const I *)row_ptr, n_nodes, keys.as<uint8_t>());
  });
  TC_HIP(hipGetLastError());
  return bucket_sort_u8(keys.as<uint8_t>(), n_nodes, kDegBuckets, perm, bucket_start, s);
}
