// Library-wide state, error reporting and device-wide scan / compaction.
#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <vector>

#include "tc_common.hpp"

namespace tc {

namespace {
thread_local char g_msg[1024] = "";
thread_local int g_code = TC_OK;
std::mutex g_init_mu;
bool g_inited[64] = {};
int g_cus[64] = {};
}  // namespace

namespace {
std::mutex g_timing_mu;
bool g_timing = false;
struct TimedLaunch {
  int stage;
  hipEvent_t a, b;
};
std::vector<TimedLaunch> g_timed;
}  // namespace

Arena *&current_arena() {
  thread_local Arena *a = nullptr;
  return a;
}

// ---------------------------------------------------------------------------
// Per-stream workspaces
// ---------------------------------------------------------------------------
namespace {
struct Chunk {
  char *base;
  size_t cap;
};
struct Workspace {
  std::vector<Chunk> chunks;
  size_t chunk = 0, off = 0;  // bump position of the open frame
};
std::mutex g_ws_mu;
std::vector<std::pair<std::pair<int, hipStream_t>, Workspace *>> g_ws;
thread_local Workspace *t_ws = nullptr;  // workspace of this thread's open frame

Workspace *workspace_for(hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto &e : g_ws)
    if (e.first.first == dev && e.first.second == s) return e.second;
  Workspace *w = new Workspace();
  g_ws.push_back({{dev, s}, w});
  return w;
}
}  // namespace

Frame::Frame(hipStream_t s) {
  if (t_ws) return;  // nested: keep carving the open frame
  Workspace *w = workspace_for(s);
  if (!w) return;
  w->chunk = 0;
  w->off = 0;
  t_ws = w;
  outer = true;
}

Frame::~Frame() {
  if (outer) t_ws = nullptr;
}

int workspace_alloc(size_t bytes, void **out) {
  *out = nullptr;
  Workspace *w = t_ws;
  if (!w) return TC_OK;  // no open frame: the caller allocates from the pool
  const size_t need = arena_round(bytes);
  while (w->chunk < w->chunks.size()) {
    Chunk &c = w->chunks[w->chunk];
    if (w->off + need <= c.cap) {
      *out = c.base + w->off;
      w->off += need;
      return TC_OK;
    }
    ++w->chunk;  // later pieces of this frame continue in the next chunk
    w->off = 0;
  }
  // grow: a chunk at least twice the last one (64 MiB minimum)
  size_t cap = std::max<size_t>(need, size_t(64) << 20);
  if (!w->chunks.empty()) cap = std::max(cap, 2 * w->chunks.back().cap);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, cap);
  if (e != hipSuccess) {
    cap = need;  // retry at the size needed
    e = hipMalloc(&p, cap);
  }
  if (e != hipSuccess) {
    set_error(e == hipErrorOutOfMemory ? TC_ERR_OOM : TC_ERR_HIP,
              "workspace growth by %zu bytes failed: %s", cap, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TC_ERR_OOM : TC_ERR_HIP;
  }
  w->chunks.push_back({static_cast<char *>(p), cap});
  w->chunk = w->chunks.size() - 1;
  w->off = need;
  *out = p;
  return TC_OK;
}

int release_workspaces() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if (!g_ws.empty() && hipDeviceSynchronize() != hipSuccess) {
    set_error(TC_ERR_HIP, "device synchronisation failed");
    return TC_ERR_HIP;
  }
  for (auto &e : g_ws) {
    for (auto &c : e.second->chunks) (void)hipFree(c.base);
    delete e.second;
  }
  g_ws.clear();
  return TC_OK;
}

bool timing_on() { return g_timing && current_arena() == nullptr; }

void timing_record(int stage, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timed.push_back({stage, a, b});
}

void set_error(int code, const char *fmt, ...) {
  g_code = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
}

int last_code() { return g_code; }

int ensure_init() {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    set_error(TC_ERR_UNSUPPORTED, "device index %d out of range", dev);
    return TC_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_inited[dev]) return TC_OK;
  hipDeviceProp_t prop;
  TC_HIP(hipGetDeviceProperties(&prop, dev));
  TC_CHECK(prop.warpSize == kWave, TC_ERR_UNSUPPORTED,
           "device %d has wavefront size %d; these kernels need 64", dev, prop.warpSize);
  g_cus[dev] = prop.multiProcessorCount;
  hipMemPool_t pool;
  TC_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t keep = UINT64_MAX;
  TC_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  g_inited[dev] = true;
  return TC_OK;
}

int num_cus() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || g_cus[dev] <= 0) return 256;
  return g_cus[dev];
}

// ---------------------------------------------------------------------------
// Reduce-then-scan.  A tile is kScanBlock threads x kScanItems items, each
// thread owning kScanItems consecutive elements.
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *total) {
  __shared__ int64_t wsum[kScanBlock / kWave];
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  int64_t inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  int64_t base = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / kWave; ++i) {
    if (i < w) base += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  if (total) *total = all;
  return base + inc - v;
}

template <class T>
__global__ void __launch_bounds__(kScanBlock)
k_tile_sums(const T *__restrict__ in, int64_t n, const int64_t *__restrict__ n_dev,
            int64_t *__restrict__ sums) {
  if (n_dev) n = min(n, *n_dev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += (int64_t)in[base + i];
  int64_t tot;
  (void)block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T>
__global__ void __launch_bounds__(kScanBlock)
k_tile_scan(const T *__restrict__ in, int64_t n, const int64_t *__restrict__ n_dev,
            const int64_t *__restrict__ offs, int64_t *__restrict__ out) {
  // with n_dev only the first m = min(n, *n_dev) inputs count; out[0..m] is
  // written (out[m] = the total) and out[n] = the total
  const int64_t m = n_dev ? min(n, *n_dev) : n;
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t vals[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = (base + i < m) ? (int64_t)in[base + i] : 0;
    s += vals[i];
  }
  int64_t tot;
  int64_t run = block_exclusive_scan(s, &tot) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i <= m && base + i < n) out[base + i] = run;
    run += vals[i];
  }
  // out has n + 1 entries: the grand total goes to out[n].
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) out[n] = run;
}

__global__ void k_zero_i64(int64_t *p) { *p = 0; }

// Scratch bytes one exclusive_scan of n elements carves (the allocation
// sequence below, every piece rounded as the arena rounds it).
size_t scan_scratch_bytes(int64_t n) {
  if (n <= 0) return 0;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) return 0;
  return arena_round(sizeof(int64_t) * tiles) + arena_round(sizeof(int64_t) * (tiles + 1)) +
         scan_scratch_bytes(tiles);
}

template <class T>
int exclusive_scan(const T *in, int64_t *out, int64_t n, int64_t * /*unused*/, hipStream_t s,
                   const int64_t *n_dev) {
  if (n <= 0) {
    // a kernel, not a memset node, so that captured graphs replay it correctly
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(1), 0, s, out);
    TC_HIP(hipGetLastError());
    return TC_OK;
  }
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kScanBlock), 0, s, in, n, n_dev,
                       (const int64_t *)nullptr, out);
    TC_HIP(hipGetLastError());
    return TC_OK;
  }
  Scratch sums, offs;
  TC_TRY(sums.alloc(sizeof(int64_t) * tiles, s));
  TC_TRY(offs.alloc(sizeof(int64_t) * (tiles + 1), s));
  hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, in, n, n_dev,
                     sums.as<int64_t>());
  TC_HIP(hipGetLastError());
  TC_TRY(exclusive_scan<int64_t>(sums.as<int64_t>(), offs.as<int64_t>(), tiles, nullptr, s));
  hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, in, n, n_dev,
                     (const int64_t *)offs.as<int64_t>(), out);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

template int exclusive_scan<int32_t>(const int32_t *, int64_t *, int64_t, int64_t *, hipStream_t,
                                   const int64_t *);
template int exclusive_scan<int64_t>(const int64_t *, int64_t *, int64_t, int64_t *, hipStream_t,
                                   const int64_t *);
template int exclusive_scan<uint8_t>(const uint8_t *, int64_t *, int64_t, int64_t *, hipStream_t,
                                   const int64_t *);
template int exclusive_scan<uint32_t>(const uint32_t *, int64_t *, int64_t, int64_t *, hipStream_t,
                                   const int64_t *);

__global__ void k_copy_i64(const int64_t *src, int64_t *dst) { *dst = *src; }

// ---------------------------------------------------------------------------
// Stable bucket sort by uint8 key.  Tiles of 64 indices (one wavefront); the
// per-(bucket, tile) counts are laid out bucket-major so one exclusive scan
// gives every output offset; ranks inside a tile come from ballot masks.
// ---------------------------------------------------------------------------
__global__ void k_bucket_count(const uint8_t *__restrict__ key, int64_t n, int64_t ntiles,
                               int nb, int32_t *__restrict__ counts) {
  const int64_t tile = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (tile >= ntiles) return;
  const int lane = lane_id();
  const int64_t i = tile * kWave + lane;
  const int b = (i < n) ? (int)key[i] : 255;
  for (int k = 0; k < nb; ++k) {
    const uint64_t m = __ballot(b == k);
    if (lane == 0) counts[(int64_t)k * ntiles + tile] = __popcll(m);
  }
}

__global__ void k_bucket_scatter(const uint8_t *__restrict__ key, int64_t n, int64_t ntiles,
                                 int nb, const int64_t *__restrict__ offs,
                                 int32_t *__restrict__ perm) {
  const int64_t tile = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (tile >= ntiles) return;
  const int lane = lane_id();
  const int64_t i = tile * kWave + lane;
  const int b = (i < n) ? (int)key[i] : 255;
  uint64_t peers = ~0ull;
#pragma unroll
  for (int bit = 0; bit < 8; ++bit) {
    const uint64_t m = __ballot((b >> bit) & 1);
    peers &= ((b >> bit) & 1) ? m : ~m;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (b < nb) perm[offs[(int64_t)b * ntiles + tile] + __popcll(peers & lt)] = (int32_t)i;
}

__global__ void k_bucket_starts(const int64_t *__restrict__ offs, int64_t ntiles, int nb,
                                int64_t *__restrict__ starts) {
  const int k = threadIdx.x;
  if (k <= nb) starts[k] = offs[(int64_t)k * ntiles];
}

size_t bucket_sort_scratch_bytes(int64_t n, int nb) {
  const int64_t ntiles = std::max<int64_t>((n + kWave - 1) / kWave, 1);
  return arena_round(sizeof(int32_t) * ntiles * nb) +
         arena_round(sizeof(int64_t) * (ntiles * nb + 1)) + scan_scratch_bytes(ntiles * nb);
}

int bucket_sort_u8(const uint8_t *key, int64_t n, int nb, int32_t *perm, int64_t *starts,
                   hipStream_t s) {
  if (nb < 1 || nb > 64) {
    set_error(TC_ERR_INVALID, "bucket count %d out of range", nb);
    return TC_ERR_INVALID;
  }
  const int64_t ntiles = std::max<int64_t>((n + kWave - 1) / kWave, 1);
  Scratch counts, offs;
  TC_TRY(counts.alloc(sizeof(int32_t) * ntiles * nb, s));
  TC_TRY(offs.alloc(sizeof(int64_t) * (ntiles * nb + 1), s));
  const unsigned g = grid_for(ntiles * kWave);
  hipLaunchKernelGGL(k_bucket_count, dim3(g), dim3(kBlock), 0, s, key, n, ntiles, nb,
                     counts.as<int32_t>());
  TC_HIP(hipGetLastError());
  TC_TRY(exclusive_scan<int32_t>(counts.as<int32_t>(), offs.as<int64_t>(), ntiles * nb, nullptr,
                                 s));
  hipLaunchKernelGGL(k_bucket_scatter, dim3(g), dim3(kBlock), 0, s, key, n, ntiles, nb,
                     (const int64_t *)offs.as<int64_t>(), perm);
  hipLaunchKernelGGL(k_bucket_starts, dim3(1), dim3(128), 0, s,
                     (const int64_t *)offs.as<int64_t>(), ntiles, nb, starts);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

// ---------------------------------------------------------------------------
// Stream compaction by flag.  A wavefront owns a chunk of 1024 flags, 16
// consecutive flags per lane read as one 16-byte load when the array is
// 16-byte aligned.  Pass 1 counts the nonzero flags of every chunk; an
// exclusive scan of the chunk counts gives each chunk its first output slot;
// pass 2 re-reads the flags and writes the indices in increasing order (lane
// offsets from a wave scan).  Traffic: two reads of the flags plus the index
// writes, and a scan over n / 1024 counts.
// ---------------------------------------------------------------------------
constexpr int kCompactPerLane = 16;
constexpr int kCompactChunk = kWave * kCompactPerLane;

// 16 flags of lane `lane` of chunk `chunk` as four little-endian words
// (bytes past n read as 0)
template <bool kVec>
__device__ __forceinline__ uint4 load_flags16(const uint8_t *__restrict__ flag, int64_t n,
                                              int64_t first) {
  if (kVec && first + kCompactPerLane <= n)
    return *reinterpret_cast<const uint4 *>(flag + first);
  uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < kCompactPerLane; ++k)
    if (first + k < n) w[k >> 2] |= (uint32_t)flag[first + k] << (8 * (k & 3));
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// 0x80 in each byte of the result whose byte in w is nonzero
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w) {
  return (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
}

__device__ __forceinline__ int count_flags16(uint4 f) {
  return __popc(nonzero_bytes(f.x)) + __popc(nonzero_bytes(f.y)) + __popc(nonzero_bytes(f.z)) +
         __popc(nonzero_bytes(f.w));
}

template <bool kVec>
__global__ void __launch_bounds__(kBlock)
k_compact_counts(const uint8_t *__restrict__ flag, int64_t n, int64_t nchunks,
                 int32_t *__restrict__ cnt) {
  const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (chunk >= nchunks) return;  // whole wavefronts leave together
  const int lane = lane_id();
  const int c = count_flags16(
      load_flags16<kVec>(flag, n, chunk * kCompactChunk + (int64_t)lane * kCompactPerLane));
  const int tot = wave_inclusive_scan(c);
  if (lane == kWave - 1) cnt[chunk] = tot;
}

template <bool kVec>
__global__ void __launch_bounds__(kBlock)
k_compact_write(const uint8_t *__restrict__ flag, int64_t n, int64_t nchunks,
                const int64_t *__restrict__ offs, int64_t max_out, int32_t *__restrict__ idx) {
  const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (chunk >= nchunks) return;
  const int lane = lane_id();
  const int64_t first = chunk * kCompactChunk + (int64_t)lane * kCompactPerLane;
  const uint4 f = load_flags16<kVec>(flag, n, first);
  const int c = count_flags16(f);
  int64_t o = offs[chunk] + (wave_inclusive_scan(c) - c);
  const uint32_t w[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
  for (int k = 0; k < kCompactPerLane; ++k)
    if ((w[k >> 2] >> (8 * (k & 3))) & 0xFFu) {
      if (o < max_out) idx[o] = (int32_t)(first + k);
      ++o;
    }
}

// *dst = min(*src, cap)
__global__ void k_copy_min_i64(const int64_t *src, int64_t *dst, int64_t cap) {
  *dst = *src < cap ? *src : cap;
}

size_t compact_scratch_bytes(int64_t n) {
  const int64_t nchunks = (n + kCompactChunk - 1) / kCompactChunk;
  return arena_round(sizeof(int32_t) * std::max<int64_t>(nchunks, 1)) +
         arena_round(sizeof(int64_t) * (nchunks + 1)) + scan_scratch_bytes(nchunks);
}

int compact_indices(const uint8_t *flag, int64_t n, int32_t *idx_out, int64_t *d_count,
                    hipStream_t s, int64_t max_out) {
  if (max_out < 0 || max_out > n) max_out = n;
  const int64_t nchunks = (n + kCompactChunk - 1) / kCompactChunk;
  Scratch cnt, offs;
  TC_TRY(cnt.alloc(sizeof(int32_t) * std::max<int64_t>(nchunks, 1), s));
  TC_TRY(offs.alloc(sizeof(int64_t) * (nchunks + 1), s));
  const bool vec = ((uintptr_t)flag & 15) == 0;
  const unsigned g = (unsigned)((nchunks * kWave + kBlock - 1) / kBlock);
  if (nchunks > 0) {
    if (vec)
      hipLaunchKernelGGL(k_compact_counts<true>, dim3(g), dim3(kBlock), 0, s, flag, n, nchunks,
                         cnt.as<int32_t>());
    else
      hipLaunchKernelGGL(k_compact_counts<false>, dim3(g), dim3(kBlock), 0, s, flag, n, nchunks,
                         cnt.as<int32_t>());
    TC_HIP(hipGetLastError());
  }
  // offs[nchunks] = number of flagged indices (0 when n == 0)
  TC_TRY(exclusive_scan<int32_t>(cnt.as<int32_t>(), offs.as<int64_t>(), nchunks, nullptr, s));
  if (nchunks > 0) {
    if (vec)
      hipLaunchKernelGGL(k_compact_write<true>, dim3(g), dim3(kBlock), 0, s, flag, n, nchunks,
                         (const int64_t *)offs.as<int64_t>(), max_out, idx_out);
    else
      hipLaunchKernelGGL(k_compact_write<false>, dim3(g), dim3(kBlock), 0, s, flag, n, nchunks,
                         (const int64_t *)offs.as<int64_t>(), max_out, idx_out);
    TC_HIP(hipGetLastError());
  }
  if (d_count) {
    hipLaunchKernelGGL(k_copy_min_i64, dim3(1), dim3(1), 0, s, offs.as<int64_t>() + nchunks,
                       d_count, max_out);
    TC_HIP(hipGetLastError());
  }
  return TC_OK;
}

}  // namespace tc

// ---------------------------------------------------------------------------
// extern "C" library entry points
// ---------------------------------------------------------------------------
extern "C" const char *tc_last_error(void) { return tc::g_msg; }

extern "C" int tc_abi_version(void) { return TC_ABI_VERSION; }

extern "C" int tc_release_workspaces(void) { return tc::release_workspaces(); }

extern "C" int tc_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(tc::g_timing_mu);
  tc::g_timing = on != 0;
  return TC_OK;
}

extern "C" int tc_timing_read(double *ms, int64_t *launches) {
  std::vector<tc::TimedLaunch> log;
  {
    std::lock_guard<std::mutex> lk(tc::g_timing_mu);
    log.swap(tc::g_timed);
  }
  for (int i = 0; i < TC_N_STAGES; ++i) {
    if (ms) ms[i] = 0.0;
    if (launches) launches[i] = 0;
  }
  int rc = TC_OK;
  for (auto &t : log) {
    float x = 0.f;
    if (hipEventSynchronize(t.b) != hipSuccess || hipEventElapsedTime(&x, t.a, t.b) != hipSuccess) {
      tc::set_error(TC_ERR_HIP, "event timing failed");
      rc = TC_ERR_HIP;
    } else if (t.stage >= 0 && t.stage < TC_N_STAGES) {
      if (ms) ms[t.stage] += x;
      if (launches) launches[t.stage] += 1;
    }
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  return rc;
}

extern "C" int tc_device_info(int *n_cu, int *wave_size, int64_t *lds_per_block,
                              int64_t *total_mem, char *name, int name_len) {
  TC_TRY(tc::ensure_init());
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  TC_HIP(hipGetDeviceProperties(&prop, dev));
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (wave_size) *wave_size = prop.warpSize;
  if (lds_per_block) *lds_per_block = (int64_t)prop.sharedMemPerBlock;
  if (total_mem) *total_mem = (int64_t)prop.totalGlobalMem;
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  return TC_OK;
}

extern "C" int tc_compact_flags(const uint8_t *flags, int64_t n, int32_t *idx, int64_t *count,
                                void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n >= 0 && n < INT32_MAX && count && (n == 0 || (flags && idx)), TC_ERR_INVALID,
           "invalid arguments to tc_compact_flags");
  TC_TRY(ensure_init());
  return compact_indices(flags, n, idx, count, (hipStream_t)stream);
}
