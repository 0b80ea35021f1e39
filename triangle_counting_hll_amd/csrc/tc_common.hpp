// Shared helpers for the gfx950 triangle-counting kernels: status handling,
// wave64 primitives, stream-ordered scratch allocation and a device-wide scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "tchll.h"

namespace tc {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // default workgroup size (4 wavefronts)

// ---------------------------------------------------------------------------
// Status handling: every extern "C" entry point returns a TC_* code and leaves
// a message retrievable through tc_last_error().
// ---------------------------------------------------------------------------
void set_error(int code, const char *fmt, ...);
int last_code();

#define TC_HIP(expr)                                                            \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::tc::set_error(TC_ERR_HIP, "%s failed: %s (%s:%d)", #expr,              \
                      hipGetErrorString(_e), __FILE__, __LINE__);               \
      return TC_ERR_HIP;                                                        \
    }                                                                           \
  } while (0)

#define TC_CHECK(cond, code, ...)                                               \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::tc::set_error((code), __VA_ARGS__);                                     \
      return (code);                                                            \
    }                                                                           \
  } while (0)

#define TC_TRY(expr)                                                            \
  do {                                                                          \
    int _rc = (expr);                                                           \
    if (_rc != TC_OK) return _rc;                                               \
  } while (0)

// Workspace arena for graph capture.  While an arena is installed on the
// calling thread, Scratch::alloc either measures (allocates normally and adds
// the rounded size to `total`) or carves consecutive pieces of `base`
// (no allocation, no free), so a pipeline whose allocation sequence depends
// only on its sizes can be replayed from one preallocated workspace.
struct Arena {
  bool carve = false;
  char *base = nullptr;
  size_t cap = 0;
  size_t total = 0;
};
Arena *&current_arena();
inline size_t arena_round(size_t b) { return (b + 255) & ~size_t(255); }

// Per-stream workspaces.  An entry point that allocates scratch opens a
// Frame on its stream; while the frame is open, Scratch::alloc carves
// consecutive pieces of that stream's persistent workspace (chunks that grow
// on first use and are kept), and the next call on the same stream starts
// again at the beginning: the stream orders the previous call's kernels
// before the next call's, so no allocation or free happens per call.  Calls
// that overlap on one stream from several host threads are not supported.
int workspace_alloc(size_t bytes, void **out);
int release_workspaces();
struct Frame {
  bool outer = false;
  explicit Frame(hipStream_t s);
  ~Frame();
  Frame(const Frame &) = delete;
  Frame &operator=(const Frame &) = delete;
};

// Scratch buffer: a piece of the graph-capture arena when one is installed,
// else a piece of the current frame's workspace, else a stream-ordered
// allocation from the device's default memory pool.
struct Scratch {
  void *ptr = nullptr;
  hipStream_t stream = nullptr;
  bool owned = false;
  Scratch() = default;
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
  ~Scratch() {
    if (ptr && owned) (void)hipFreeAsync(ptr, stream);
  }
  int alloc(size_t bytes, hipStream_t s) {
    stream = s;
    if (bytes == 0) bytes = 256;
    Arena *ar = current_arena();
    if (ar && ar->carve) {
      const size_t need = arena_round(bytes);
      if (ar->total + need > ar->cap) {
        set_error(TC_ERR_OOM, "graph workspace exhausted (%zu of %zu bytes)", ar->total + need,
                  ar->cap);
        return TC_ERR_OOM;
      }
      ptr = ar->base + ar->total;
      ar->total += need;
      owned = false;
      return TC_OK;
    }
    if (ar) ar->total += arena_round(bytes);
    if (!ar) {
      void *w = nullptr;
      const int rc = workspace_alloc(bytes, &w);
      if (rc != TC_OK) return rc;
      if (w) {
        ptr = w;
        owned = false;
        return TC_OK;
      }
    }
    owned = true;
    hipError_t e = hipMallocAsync(&ptr, bytes, s);
    if (e != hipSuccess) {
      ptr = nullptr;
      set_error(e == hipErrorOutOfMemory ? TC_ERR_OOM : TC_ERR_HIP,
                "hipMallocAsync(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? TC_ERR_OOM : TC_ERR_HIP;
    }
    return TC_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Stage timing (see tc_timing_enable): RAII pair of events around a launch.
// Off while a pipeline is being captured into a graph.
bool timing_on();
void timing_record(int stage, hipEvent_t a, hipEvent_t b);
struct StageTimer {
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(int st, hipStream_t str) : stage(st), s(str) {
    if (timing_on() && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, s);
    else
      a = b = nullptr;
  }
  ~StageTimer() {
    if (a && b) {
      (void)hipEventRecord(b, s);
      timing_record(stage, a, b);
    }
  }
};

int ensure_init();   // once per process: device pool threshold, device facts
int num_cus();       // compute units of the current device

inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Wave64 primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Completes this wavefront's outstanding LDS traffic and keeps the compiler
// from moving memory operations across the point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <class T> __device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T t = __shfl_up(v, off, kWave);
    if (lane >= off) v += t;
  }
  return v;
}

template <class T> __device__ __forceinline__ T wave_broadcast(T v, int src) {
  return __shfl(v, src, kWave);
}

// Row pointer accessor for int32 or int64 CSR offsets.
template <class I> struct Rows {
  const I *p;
  __device__ __forceinline__ int64_t begin(int64_t u) const { return (int64_t)p[u]; }
  __device__ __forceinline__ int64_t end(int64_t u) const { return (int64_t)p[u + 1]; }
  __device__ __forceinline__ int64_t deg(int64_t u) const { return (int64_t)p[u + 1] - (int64_t)p[u]; }
};

// Dispatch a functor templated on the row-pointer width.
#define TC_DISPATCH_ROWS(is64, ROWT, ...)                                        \
  do {                                                                          \
    if (is64) {                                                                 \
      using ROWT = int64_t;                                                     \
      __VA_ARGS__;                                                              \
    } else {                                                                    \
      using ROWT = int32_t;                                                     \
      __VA_ARGS__;                                                              \
    }                                                                           \
  } while (0)

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (reduce-then-scan), int64 output.
// out[i] = sum_{j<i} in[i] for i <= n (out has n + 1 entries).  With n_dev
// (device), only the first m = min(n, *n_dev) inputs are read: out[0..m] and
// out[n] are written (both ends hold the total), the rest is left alone.
// ---------------------------------------------------------------------------
template <class T>
int exclusive_scan(const T *in, int64_t *out, int64_t n, int64_t *d_total, hipStream_t s,
                   const int64_t *n_dev = nullptr);

// Stable bucket sort of 0..n-1 by a small key (key[i] < nbuckets <= 64, keys
// equal to nbuckets or more are dropped).  perm receives the indices grouped
// by ascending key, ascending index inside a key; starts[k] (k = 0..nbuckets,
// device int64) is the offset of bucket k in perm, starts[nbuckets] the total.
int bucket_sort_u8(const uint8_t *key, int64_t n, int nbuckets, int32_t *perm, int64_t *starts,
                   hipStream_t s);

// Stable LSD radix sort of 64-bit words by their bits [bit_lo, bit_hi), 8 bits
// per pass.  `a` holds the input; `b` is a buffer of the same size.  On return
// *result points at whichever of the two holds the sorted words.  With `drop`
// (host pointer to a key value), the first pass leaves out every word equal to
// *drop and the sorted output holds only the others; then n_dev (device) must
// hold their number, which the later passes read (n sizes the launches).
// With n_dev and no drop, every pass reads only the first *n_dev words.
// With run_flags (uint8, n entries), the last pass also writes
// run_flags[i] = 1 iff sorted word i starts a run of equal bits [bit_lo, 64)
// (i == 0 or those bits differ from word i - 1), 0 otherwise, for every i
// below the number of sorted words.  With bit_hi < 64 the words are ordered by
// bits [bit_lo, bit_hi) only, so one value of bits [bit_lo, 64) can form
// several runs unless the caller's input order keeps its words together.
int radix_sort_u64(uint64_t *a, uint64_t *b, int64_t n, int bit_lo, int bit_hi,
                   uint64_t **result, hipStream_t s, const uint64_t *drop = nullptr,
                   const unsigned long long *n_dev = nullptr, uint8_t *run_flags = nullptr);

// Stream compaction by flag: idx_out receives the indices i with flag[i] != 0 in
// increasing order; *d_count (device) receives their number.  With max_out
// >= 0, only the first max_out of them are written and *d_count is at most
// max_out (idx_out then needs max_out entries, not n).
int compact_indices(const uint8_t *flag, int64_t n, int32_t *idx_out, int64_t *d_count,
                    hipStream_t s, int64_t max_out = -1);

// Scratch bytes each helper above carves from an arena for a problem of the
// given size (the sum of its Scratch pieces, rounded as arena_round does).
size_t scan_scratch_bytes(int64_t n);
size_t bucket_sort_scratch_bytes(int64_t n, int nbuckets);
size_t radix_sort_scratch_bytes(int64_t n, int bit_lo, int bit_hi, bool run_flags = false);
size_t compact_scratch_bytes(int64_t n);

}  // namespace tc
