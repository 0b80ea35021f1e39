// (b) Exact per-edge set intersection on gfx950.
//
// T = sum over oriented edges (u -> v) of |N+(u) ∩ N+(v)|, both rows sorted.
//
// Pipeline (stream-ordered; small inputs read one counter back, see
// exact_sorted_plan):
//   1. classify: an edge whose two rows together hold <= light_max ids is
//      merged by one lane right away.  Every other edge is assigned to its
//      "center" = the endpoint with the longer row (ties: v); the other
//      endpoint's row will be streamed against the center's row.  Each heavy
//      edge becomes one packed word: the center in the top bits, below it the
//      other's row (start, length), or the other's id when the row fields do
//      not fit (PackLayout).
//   2. LSD radix sort of the packed words on the center bits, then the runs
//      of equal centers (run starts flagged and compacted) -> for each center
//      the list of its "others", in edge order.  (The atomic plan, kept as an
//      option, groups them with a counting sort instead.)
//   3. work items = (run, chunk of <= `chunk` others), longest center row
//      first (bucket sort on the row length's bit length).
//   4. persistent heavy kernels: a wavefront (or a workgroup, for long center
//      rows) pulls an item, builds an LDS open-addressing hash set of
//      N+(center), then streams the concatenation of the others' rows with
//      all 64 lanes (flattened across rows, so short rows do not idle lanes),
//      16 ids in flight per lane, and probes each id.
// Streaming the shorter row into a hash of the longer one costs
// sum_e min(d+u, d+v) probes, against sum_e (d+u + d+v) for a merge.
#include <algorithm>
#include <mutex>

#include "tc_common.hpp"
#include "tc_exact.hpp"
#include "tc_hll.hpp"
#include "tc_router.hpp"

namespace tc {

constexpr int kTblInts = 2048;    // per-wave LDS hash table: 512 buckets x 4 slots (8 KiB)
constexpr int kTblMaxDeg = kWaveTableMaxDeg;  // longer center rows: one table per workgroup
constexpr int kBlkMinBits = 10;   // default: d+ in [512, 4095] take the workgroup's four wave tables
constexpr int kBlkTblInts = 4 * kTblInts;  // joined into one (2048 buckets, 32 KiB)
constexpr int kBigMinBits = 13;   // d+ >= 4096: the 128 KiB workgroup kernel
constexpr int64_t kEarlyExitMaxEdges = int64_t(1) << 24;  // see exact_sorted_plan
constexpr int kBigTblInts = 32768;  // block-shared table: 8192 buckets x 4 slots (128 KiB)
constexpr int kBigMaxDeg = kBigTblInts / 2 - 1;  // longer rows: global binary search
constexpr int kDefaultUnroll = 16;  // streamed ids in flight per lane
constexpr int kWavesPerBlock = kBlock / kWave;

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// CSR mode (no edge list): eu[k] = the row of position k, for the positions
// k < E that the rows reach.  The caller's n_edges (E) should equal
// row_ptr[n]; positions past row_ptr[n] are left unwritten and the edge
// passes skip them (edge_limit), positions the rows reach past E are never
// written.
template <class I>
__global__ void k_expand_rows(const I *__restrict__ rp, int64_t n, int64_t E,
                              int32_t *__restrict__ eu) {
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t u = wid; u < n; u += nw) {
    const int64_t hi = min((int64_t)rp[u + 1], E);
    for (int64_t k = (int64_t)rp[u] + lane; k < hi; k += kWave) eu[k] = (int32_t)u;
  }
}

// Number of edges an edge pass may read: E, or in CSR mode min(E, row_ptr[n])
// (row_ptr[n] measured on the device, stats[1] of k_row_stats).
__device__ __forceinline__ int64_t edge_limit(int64_t E, const unsigned long long *rows_total) {
  return rows_total ? min(E, (int64_t)*rows_total) : E;
}

// last[v] = the largest id of row v (its last; INT32_MIN for an empty row).
template <class I>
__global__ void k_last_ids(Rows<I> R, const int32_t *__restrict__ col, int64_t n,
                           int32_t *__restrict__ last) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = R.begin(v), e = R.end(v);
    last[v] = e > b ? col[e - 1] : INT32_MIN;
  }
}

// info[v] = {row start (low and high words), length, largest id (INT32_MIN
// for an empty row)}: what the trimming edge pass needs of an endpoint, in
// one aligned 16-byte entry.
template <class I>
__global__ void k_row_info(Rows<I> R, const int32_t *__restrict__ col, int64_t n,
                           int4 *__restrict__ info) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = R.begin(v), e = R.end(v);
    info[v] = make_int4((int32_t)(uint32_t)b, (int32_t)(b >> 32), (int32_t)(e - b),
                        e > b ? col[e - 1] : INT32_MIN);
  }
}

// Number of ids <= hi in the sorted row col[st, st + len) whose last id is
// above hi.  An 8-ary search: each round reads the last ids of the first
// seven eighths of the range together (one memory round trip instead of
// three for a binary search) and keeps the eighth that holds the boundary.
__device__ __forceinline__ int64_t count_at_most(const int32_t *__restrict__ col, int64_t st,
                                                 int64_t len, int32_t hi) {
  // the answer lies in [lo, lo + m]: ids before lo are <= hi, ids from
  // lo + m on are above it (the last one is)
  int64_t lo = 0, m = len - 1;
  while (m >= 8) {
    const int64_t q = m >> 3;
    int k = 0;  // eighths whose last id is <= hi (a prefix of them)
#pragma unroll
    for (int j = 1; j < 8; ++j) k += col[st + lo + j * q - 1] <= hi ? 1 : 0;
    lo += k * q;
    m = k == 7 ? m - 7 * q : q - 1;
  }
  while (m > 0) {
    const int64_t half = m >> 1;
    if (col[st + lo + half] <= hi) {
      lo += half + 1;
      m -= half + 1;
    } else {
      m = half;
    }
  }
  return lo;
}

// count_at_most for a row whose answer is probably near a known position:
// the answer A (the first position of the row col[ub, ub + du) holding an
// id above hi; the last one does) is bracketed by [a, b].  The ids of a
// window at row positions inside [a, b] are sorted, so those <= hi are a
// prefix of them: c of nv give a = (first position) + c when c > 0 and
// b = (first position) + c when c < nv.  First the kTrimWin ids of col from
// qs on are read with 16-byte loads (qs is clamped so that the reads stay
// inside col[0, total)); when the boundary is not among them, the next
// window on the side that holds it; then the 8-ary search of what is left of
// [a, b].  In the CSR's own edge order the edge u -> v at index e has v at
// col[e], and on graphs whose ids follow the degrees the answer mostly lies
// within a few ids of it, so the caller centres the first window on e + 1.
constexpr int kTrimWin = 16;

// The first window of the search for the edge at index e: its start in col
// (centred on e + 1, clamped inside col) and its kTrimWin ids.
struct TrimWindow {
  int64_t qs;
  int4 w[kTrimWin / 4];
};
__device__ __forceinline__ void trim_window_load(const int32_t *__restrict__ col, int64_t e,
                                                 int64_t total, TrimWindow &T) {
  T.qs = max((int64_t)0, min(e + 1 - kTrimWin / 2, total - kTrimWin));
#pragma unroll
  for (int i = 0; i < kTrimWin / 4; ++i) __builtin_memcpy(&T.w[i], col + T.qs + 4 * i, 16);
}

// Narrows [a, b] by the window of ids w starting at col position qs (row
// positions qs - ub on); returns -1, 0 or 1: look left, stop, look right.
__device__ __forceinline__ int trim_window_step(const int4 (&w)[kTrimWin / 4], int64_t qs,
                                                int64_t ub, int du, int32_t hi, int &a, int &b) {
  // row position of the window's first id, clamped to [-kTrimWin, du]: a
  // window wholly outside the row keeps its side of it
  const int p0 = (int)max((int64_t)-kTrimWin, min(qs - ub, (int64_t)du));
  const int pf = max(p0, a), pl = min(p0 + kTrimWin - 1, b);
  const int nv = pl - pf + 1;  // window positions inside [a, b]
  if (nv <= 0) return p0 + kTrimWin <= a ? 1 : (p0 > b ? -1 : 0);
  const uint32_t span = (uint32_t)(b - a);
  int c = 0;
#pragma unroll
  for (int i = 0; i < kTrimWin / 4; ++i) {
    const int32_t x[4] = {w[i].x, w[i].y, w[i].z, w[i].w};
#pragma unroll
    for (int j = 0; j < 4; ++j) c += ((uint32_t)(p0 + 4 * i + j - a) <= span && x[j] <= hi) ? 1 : 0;
  }
  if (c > 0) a = pf + c;
  if (c < nv) b = pf + c;
  return c == nv ? 1 : (c == 0 ? -1 : 0);
}

// count_at_most of the row col[ub, ub + du) given its first window T (read
// by trim_window_load for the edge's index e): then, while the boundary is
// not found, the next window on the side that holds it, then the 8-ary
// search of what is left of [a, b].  total >= kTrimWin.
__device__ __forceinline__ int64_t count_at_most_near(const int32_t *__restrict__ col, int64_t ub,
                                                      int du, int32_t hi, const TrimWindow &T,
                                                      int64_t total) {
  int a = 0, b = du - 1;
  int dir = trim_window_step(T.w, T.qs, ub, du, hi, a, b);
  if (dir != 0 && a < b) {
    const int64_t qs = max((int64_t)0, min(T.qs + dir * kTrimWin, total - kTrimWin));
    int4 w[kTrimWin / 4];
#pragma unroll
    for (int i = 0; i < kTrimWin / 4; ++i) __builtin_memcpy(&w[i], col + qs + 4 * i, 16);
    trim_window_step(w, qs, ub, du, hi, a, b);
  }
  return a + (a < b ? count_at_most(col, ub + a, b - a + 1, hi) : 0);
}

__device__ __forceinline__ uint32_t merge_count(const int32_t *__restrict__ col, int64_t a0,
                                                int64_t a1, int64_t b0, int64_t b1) {
  uint32_t c = 0;
  if (a0 >= a1 || b0 >= b1) return 0;
  int x = col[a0], y = col[b0];
  while (true) {
    if (x == y) {
      ++c;
      if (++a0 >= a1 || ++b0 >= b1) break;
      x = col[a0];
      y = col[b0];
    } else if (x < y) {
      if (++a0 >= a1) break;
      x = col[a0];
    } else {
      if (++b0 >= b1) break;
      y = col[b0];
    }
  }
  return c;
}

// Light edges of one wavefront round, each with rows of at most 32 ids.  When
// a round holds few light edges, the wavefront intersects them one after the
// other: one load puts row a in lanes [0, la) and row b in lanes [32, 32 + lb),
// then one ballot per id of a counts its match in b.  With many light edges in
// the round, each lane intersects its own from registers (lane_intersection;
// a serial merge, a chain of dependent loads, remains for rows too long for
// them).  Returns |a ∩ b| in the edge's lane, 0 elsewhere.
constexpr int kCoopMaxPerRound = 16;

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One light edge per lane without a chain of dependent loads: the shorter
// row (at most K ids) is loaded into registers at once, the longer one
// streamed four ids per step, and every streamed id is compared with the
// K registers (ids are non-negative; the fill values match nothing).
template <int K>
__device__ __forceinline__ uint32_t lane_intersection(const int32_t *__restrict__ col, int64_t s0,
                                                      int ds, int64_t l0, int dl) {
  int S[K];
#pragma unroll
  for (int i = 0; i < K; ++i) S[i] = (i < ds) ? col[s0 + i] : -1;
  uint32_t c = 0;
  for (int j = 0; j < dl; j += 4) {
    int y[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = (j + t < dl) ? col[l0 + j + t] : -2;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < K; ++i) c += (S[i] == y[t]) ? 1u : 0u;
  }
  return c;
}

__device__ __forceinline__ uint32_t light_intersections(const int32_t *__restrict__ col, bool coop,
                                                        int64_t a0, int la, int64_t b0, int lb) {
  uint64_t lm = __ballot(coop);
  if (!lm) return 0;
  if (__popcll(lm) > kCoopMaxPerRound) {
    // many light edges: one per lane, the register width chosen per wavefront
    int64_t s0 = a0, l0 = b0;
    int ds = la, dl = lb;
    if (la > lb) {
      s0 = b0;
      l0 = a0;
      ds = lb;
      dl = la;
    }
    if (!coop) ds = dl = 0;
    if (!__ballot(ds > 4)) return lane_intersection<4>(col, s0, ds, l0, dl);
    if (!__ballot(ds > 8)) return lane_intersection<8>(col, s0, ds, l0, dl);
    if (!__ballot(ds > 16)) return lane_intersection<16>(col, s0, ds, l0, dl);
    return coop ? merge_count(col, a0, a0 + la, b0, b0 + lb) : 0u;
  }
  const int lane = lane_id();
  constexpr int kHalf = kWave / 2;
  uint32_t mine = 0;
  // rows of the light edge of lane `src`: a in lanes [0, na), b in [32, 32 + nb)
  auto fetch = [&](int src, int &na) -> int {
    const int64_t ea = readlane_i64(a0, src), eb = readlane_i64(b0, src);
    na = __builtin_amdgcn_readlane(la, src);
    const int nb = __builtin_amdgcn_readlane(lb, src);
    int x = -2;  // matches no id
    if (lane < na) x = col[ea + lane];
    else if (lane >= kHalf && lane - kHalf < nb) x = col[eb + (lane - kHalf)];
    return x;
  };
  int src = __ffsll((long long)lm) - 1;
  lm &= lm - 1;
  int na = 0;
  int x = fetch(src, na);
  while (true) {
    // the next edge's rows are in flight while this edge's ids are compared
    const bool more = lm != 0;
    int src2 = 0, na2 = 0, x2 = -2;
    if (more) {
      src2 = __ffsll((long long)lm) - 1;
      lm &= lm - 1;
      x2 = fetch(src2, na2);
    }
    uint32_t c = 0;
    for (int i = 0; i < na; ++i) {
      const int ai = __builtin_amdgcn_readlane(x, i);
      c += (uint32_t)__popcll(__ballot(lane >= kHalf && x == ai));
    }
    if (lane == src) mine = c;
    if (!more) break;
    src = src2;
    na = na2;
    x = x2;
  }
  return mine;
}

// Edge filter shared by the exact and hybrid pipelines: mask == nullptr means
// every edge is counted; otherwise edges with mask[e] != 0 are skipped.
template <class I>
__global__ void __launch_bounds__(kBlock)
k_classify(Rows<I> R, const int32_t *__restrict__ col, const int32_t *__restrict__ eu,
           const int32_t *__restrict__ ev, int64_t E, const unsigned long long *rows_total,
           const uint8_t *__restrict__ skip, int light_max, int32_t *__restrict__ center_cnt,
           const int32_t *__restrict__ graph_of, unsigned long long *__restrict__ count) {
  uint64_t acc = 0;
  E = edge_limit(E, rows_total);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (skip && skip[e]) continue;
    const int32_t u = eu[e], v = ev[e];
    const int64_t ub = R.begin(u), ue = R.end(u), vb = R.begin(v), ve = R.end(v);
    const int64_t du = ue - ub, dv = ve - vb;
    if (du == 0 || dv == 0) continue;
    if (du + dv <= light_max) {
      uint32_t c = merge_count(col, ub, ue, vb, ve);
      if (graph_of) {
        if (c) atomicAdd(&count[graph_of[u]], (unsigned long long)c);
      } else {
        acc += c;
      }
    } else {
      atomicAdd(&center_cnt[dv >= du ? v : u], 1);
    }
  }
  if (!graph_of) {
    acc = wave_sum(acc);
    if (lane_id() == 0 && acc) atomicAdd(count, (unsigned long long)acc);
  }
}

template <class I>
__global__ void __launch_bounds__(kBlock)
k_scatter_others(Rows<I> R, const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                 int64_t E, const unsigned long long *rows_total,
                 const uint8_t *__restrict__ skip, int light_max,
                 const int64_t *__restrict__ cptr, int32_t *__restrict__ cursor,
                 int32_t *__restrict__ others) {
  E = edge_limit(E, rows_total);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (skip && skip[e]) continue;
    const int32_t u = eu[e], v = ev[e];
    const int64_t du = R.deg(u), dv = R.deg(v);
    if (du == 0 || dv == 0 || du + dv <= light_max) continue;
    const bool cv = dv >= du;
    const int32_t c = cv ? v : u, o = cv ? u : v;
    others[cptr[c] + atomicAdd(&cursor[c], 1)] = o;
  }
}

// Work items are ordered longest center row first: key = 32 - bit_length(d+c)
// for centers with heavy edges, no key (255) otherwise.
constexpr int kCenterBuckets = 32;
constexpr int kBigCenterBits = 8;  // centers with d+ >= 128 get one item per queue ticket
constexpr int kSmallBatch = 16;    // items per ticket for the shorter centers
constexpr int kBandTicket = 4;     // items per ticket with item_order (one graph)
constexpr int kBatchTicket = 8;    // items per ticket of a batch's per-XCD queues
constexpr int kXcds = 8;           // XCDs of an MI355X (one L2 each)
constexpr int kXcdStride = 16;     // u64 counters between two XCD queues (128 B)

template <class I>
__global__ void k_center_keys(Rows<I> R, const int32_t *__restrict__ center_cnt, int64_t n,
                              uint8_t *__restrict__ key) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t d = (uint64_t)R.deg(c);
    key[c] = (center_cnt[c] > 0 && d > 0) ? (uint8_t)(kCenterBuckets - (64 - __clzll((long long)d)))
                                          : (uint8_t)255;
  }
}

__global__ void k_item_counts(const int32_t *__restrict__ center_cnt,
                              const int32_t *__restrict__ perm, const int64_t *__restrict__ n_cent_p,
                              int64_t n, int chunk, int32_t *__restrict__ nitems) {
  const int64_t nc = *n_cent_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    nitems[i] = i < nc ? (center_cnt[perm[i]] + chunk - 1) / chunk : 0;
}

__global__ void k_item_fill(const int32_t *__restrict__ nitems, const int32_t *__restrict__ perm,
                            const int64_t *__restrict__ iptr, const int64_t *__restrict__ n_cent_p,
                            int2 *__restrict__ items) {
  const int64_t nc = *n_cent_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = nitems[i];
    const int c = perm[i];
    const int64_t b = iptr[i];
    for (int j = 0; j < k; ++j) items[b + j] = make_int2(c, j);
  }
}

// LDS hash set of one center row: buckets of 4 int slots read with one
// 16-byte LDS load.  A bucket's slots fill in order 0..3 and a key moves to the
// next bucket only when its home bucket is full, so a lookup stops at the
// first bucket whose last slot is empty.  Every table uses all the buckets its
// LDS holds, whatever the row length: the fewer keys per bucket, the rarer a
// full home bucket, and a wavefront skips the continuation probes of a step
// when none of its lanes met one.  Load factor <= 1/2 (kTblMaxDeg ids in a
// 512-bucket wave table).
__device__ __forceinline__ uint32_t bucket_of(uint32_t x, int shift) {
  return (x * 0x9E3779B1u) >> shift;
}

// Builds the set in `tbl` (kInts ints) with the lanes of one wavefront
// (kBlockWide false) or all threads of the workgroup (true).
template <int kInts, bool kBlockWide>
__device__ __forceinline__ void table_build(int *__restrict__ tbl, const int32_t *__restrict__ row,
                                            int dc, int &shift, uint32_t &mask) {
  static_assert(kInts >= 16 && (kInts & (kInts - 1)) == 0, "table of 2^k buckets of 4 ints");
  constexpr int NB = kInts / 4;
  const int t0 = kBlockWide ? (int)threadIdx.x : lane_id();
  const int stride = kBlockWide ? (int)blockDim.x : kWave;
  shift = 32 - __builtin_ctz((unsigned)NB);
  mask = (uint32_t)NB - 1;
  int4 *tbl4 = reinterpret_cast<int4 *>(tbl);  // the tables are int4 arrays
  for (int i = t0; i < NB; i += stride) tbl4[i] = make_int4(-1, -1, -1, -1);
  if constexpr (kBlockWide) __syncthreads(); else wave_lds_sync();
  for (int i = t0; i < dc; i += stride) {
    const int x = row[i];
    uint32_t b = bucket_of((uint32_t)x, shift);
    while (true) {
      int *sl = tbl + 4 * b;
      if (atomicCAS(sl, -1, x) == -1 || atomicCAS(sl + 1, -1, x) == -1 ||
          atomicCAS(sl + 2, -1, x) == -1 || atomicCAS(sl + 3, -1, x) == -1)
        break;
      b = (b + 1) & mask;
    }
  }
  if constexpr (kBlockWide) __syncthreads(); else wave_lds_sync();
}

// How a work item's streamed ids are tested against its center row.
constexpr int kProbeGlobal = 0;  // binary search of the row in global memory
constexpr int kProbeHash = 1;    // LDS hash set (table_build)
constexpr int kProbeBitmap = 2;  // LDS bitmap over the row's id range (bitmap_build)

// Bitmap of one center row over its id range: bit (x - lo) is set for every
// id x of the row (lo = the row's first, smallest id).  Bit `sentinel` (the
// one after the row's largest id) stays clear: a probe clamps every offset
// to it, so ids outside the range read a zero bit with no branch.  Only the
// words up to the sentinel are cleared.  Needs sentinel < kInts * 32.
template <int kInts, bool kBlockWide>
__device__ __forceinline__ void bitmap_build(int *__restrict__ tbl, const int32_t *__restrict__ row,
                                             int dc, int32_t lo, uint32_t sentinel) {
  const int t0 = kBlockWide ? (int)threadIdx.x : lane_id();
  const int stride = kBlockWide ? (int)blockDim.x : kWave;
  const int n4 = (int)((sentinel >> 7) + 1);  // int4 groups covering words 0 .. sentinel / 32
  int4 *tbl4 = reinterpret_cast<int4 *>(tbl);
  for (int i = t0; i < n4; i += stride) tbl4[i] = make_int4(0, 0, 0, 0);
  if constexpr (kBlockWide) __syncthreads(); else wave_lds_sync();
  uint32_t *words = reinterpret_cast<uint32_t *>(tbl);
  for (int i = t0; i < dc; i += stride) {
    const uint32_t off = (uint32_t)(row[i] - lo);
    atomicOr(&words[off >> 5], 1u << (off & 31));
  }
  if constexpr (kBlockWide) __syncthreads(); else wave_lds_sync();
}

// The bitmap form fits a table of kInts ints when the row's id range plus the
// sentinel bit fits its bits; *lo and *sentinel receive the bitmap's origin
// and sentinel index.  Wave-uniform (every lane reads the same two ids).
template <int kInts>
__device__ __forceinline__ bool bitmap_fits(const int32_t *__restrict__ row, int dc, int32_t *lo,
                                            uint32_t *sentinel) {
  const int32_t a = row[0], b = row[dc - 1];
  *lo = a;
  *sentinel = (uint32_t)(b - a) + 1u;
  return (uint64_t)(uint32_t)(b - a) + 2u <= (uint64_t)kInts * 32u;
}

// 0 iff x is one of the bucket's four slots (distance by xor, straight-line
// VALU work with no per-slot branches).
__device__ __forceinline__ uint32_t bucket_miss(int4 v, int x) {
  return min(min((uint32_t)(v.x ^ x), (uint32_t)(v.y ^ x)),
             min((uint32_t)(v.z ^ x), (uint32_t)(v.w ^ x)));
}

// Membership of x (not -1, the empty-slot marker) probing buckets b, b + 1, ...
// until x or a bucket whose last slot is empty.
__device__ __forceinline__ uint32_t table_probe_from(const int4 *__restrict__ tbl, int x,
                                                     uint32_t b, uint32_t mask) {
  while (true) {
    const int4 v = tbl[b];  // one ds_read_b128
    const uint32_t miss = bucket_miss(v, x);
    if (miss == 0) return 1u;
    if (v.w == -1) return 0u;
    b = (b + 1) & mask;
  }
}

__device__ __forceinline__ uint32_t global_probe(const int32_t *__restrict__ col, int64_t cb,
                                                 int dc, int x) {
  int64_t lo = cb, hi = cb + dc;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (col[mid] < x) lo = mid + 1; else hi = mid;
  }
  return (lo < cb + dc && col[lo] == x) ? 1u : 0u;
}

// Row statistics measured on the device before a count (k_row_stats):
// stats[0] = the largest forward degree, stats[1] = row_ptr[n], the number of
// ids in col[] that the rows reach, stats[2] = 1 to force 64-bit id addressing
// (the wide_offsets parameter).  Every choice below is made from them
// inside the kernels, so it needs no host synchronisation and holds for any
// memory the arrays live in.
__device__ __forceinline__ int bit_length_u64(unsigned long long x) {
  return x ? 64 - __clzll((long long)x) : 0;
}

// Narrow id addressing: every id position of the rows fits a 32-bit byte
// offset (and the caller did not ask for the wide form).
__device__ __forceinline__ bool rows_narrow(const unsigned long long *stats) {
  return stats[2] == 0 && stats[1] <= (1ull << 30);
}

// Layout of the sorted plan's packed words.  The center id fills the top
// 64 - shift bits.  Below it, either
//   row layout: the other's row start in bits [lbits, shift) and its length
//               in bits [0, lbits), so the intersection kernel reads no row
//               offsets for the streamed rows, or
//   id layout:  the other's id in the low 32 bits.
// The row layout is used when the start field (lbits = shift - its width)
// leaves at least one bit for lengths and every forward degree is below
// 2^lbits, unless force_ids.  The start field holds row_ptr[other] in
// bit_length(row_ptr[n]) bits, or, with rel_bits > 0, row_ptr[other] -
// row_ptr[center] + 2^(rel_bits - 1) in rel_bits bits: the caller guarantees
// |row_ptr[other] - row_ptr[center]| < 2^(rel_bits - 1) for every edge (a
// batch of graphs whose rows each span fewer ids than that).
struct PackLayout {
  int shift;
  int force_ids;
  const unsigned long long *stats;
  int rel_bits = 0;
  // row layout, other = u: with last_id (the largest id of each row), the
  // length field holds the number of u's ids that are <= the center row's
  // largest id (the rest cannot be in the center row), found by the edge
  // pass; else the other's whole length
  const int32_t *last_id = nullptr;
  // the same trimming with each endpoint's row start, length and last id in
  // one entry (k_row_info), for the edge pass instances with kNear
  const int4 *row_info = nullptr;
  __device__ __forceinline__ int lbits() const {
    return shift - (rel_bits > 0 ? rel_bits : max(1, bit_length_u64(stats[1])));
  }
  // start field of an other's row (row layout)
  __device__ __forceinline__ uint64_t start_field(int64_t other_start,
                                                  int64_t center_start) const {
    return rel_bits > 0
               ? (uint64_t)(other_start - center_start + (int64_t(1) << (rel_bits - 1)))
               : (uint64_t)other_start;
  }
  __device__ __forceinline__ bool rows() const {
    const int lb = lbits();
    return !force_ids && lb > 0 && stats[0] < (1ull << lb);
  }
};

// The "others" of a work item, in either planning form: a plain id array
// indexed through per-center offsets (item.x = center), or the sorted packed
// words indexed through run offsets (item.x = run).  row() gives the row of
// the k-th other; resolve() is called once per kernel before any row().
struct OthersIds {
  const int32_t *p;
  __device__ __forceinline__ void resolve() {}
  template <class I>
  __device__ __forceinline__ void row(Rows<I> R, int64_t k, int64_t, int64_t &st,
                                      int &len) const {
    const int32_t o = p[k];
    st = R.begin(o);
    len = (int)(R.end(o) - st);
  }
  __device__ __forceinline__ int32_t center(int item_x, int64_t) const { return item_x; }
};
struct OthersPacked {
  const uint64_t *p;
  PackLayout L;
  bool rows_layout = false;
  int lb = 0;
  __device__ __forceinline__ void resolve() {
    rows_layout = L.rows();
    lb = L.lbits();
  }
  // cb: row_ptr[center] of the work item (the base of relative starts)
  template <class I>
  __device__ __forceinline__ void row(Rows<I> R, int64_t k, int64_t cb, int64_t &st,
                                      int &len) const {
    const uint64_t w = p[k];
    if (rows_layout) {
      st = (int64_t)((w >> lb) & ((1ull << (L.shift - lb)) - 1));
      if (L.rel_bits > 0) st += cb - (int64_t(1) << (L.rel_bits - 1));
      len = (int)(w & ((1ull << lb) - 1));
    } else {
      const int32_t o = (int32_t)(uint32_t)w;
      st = R.begin(o);
      len = (int)(R.end(o) - st);
    }
  }
  __device__ __forceinline__ int32_t center(int, int64_t ob) const {
    return (int32_t)(p[ob] >> L.shift);
  }
};

// Position of the id at flattened offset `off + t` of col[].  With kNarrow
// (every row within the first 2^30 ids, see rows_narrow) the offset is a
// 32-bit byte offset added to the array base, one VGPR and one 32-bit add per
// id instead of 64-bit address arithmetic; `off` may be "negative" (it wraps),
// only the sum is in range.
template <bool kNarrow>
struct IdOffset {
  using T = int64_t;
  __device__ __forceinline__ static int32_t load(const int32_t *__restrict__ col, T off, int t) {
    return col[off + t];
  }
  __device__ __forceinline__ static int4 load4(const int32_t *__restrict__ col, T off, int t) {
    int4 v;
    __builtin_memcpy(&v, col + off + t, sizeof(v));
    return v;
  }
};
template <>
struct IdOffset<true> {
  using T = uint32_t;
  __device__ __forceinline__ static int32_t load(const int32_t *__restrict__ col, T off, int t) {
    const uint32_t b = (off + (uint32_t)t) << 2;
    return *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(col) + b);
  }
  // ids off + t .. off + t + 3 with one 16-byte load (4-byte aligned)
  __device__ __forceinline__ static int4 load4(const int32_t *__restrict__ col, T off, int t) {
    const uint32_t b = (off + (uint32_t)t) << 2;
    const int32_t *p = reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(col) + b);
    int4 v;
    __builtin_memcpy(&v, p, sizeof(v));
    return v;
  }
};

// The center row's set for one work item: an LDS hash table (tbl, shift,
// mask of table_build), an LDS bitmap (tbl, lo, sentinel of bitmap_build) or
// the row itself in global memory (cb, dc).
struct CenterSet {
  const int4 *tbl;
  int shift;
  uint32_t mask;
  int32_t lo;
  uint32_t sentinel;
  int64_t cb;
  int dc;
  int64_t total_ids;  // row_ptr[n]: ids of col[] the rows reach
};

// Counts the ids of xs[0:K] found in the center's set (kMode: kProbeGlobal,
// kProbeHash or kProbeBitmap).  Padding slots hold -2, which matches no id,
// no empty hash slot and no bit of a bitmap.
template <int kMode, int K>
__device__ __forceinline__ uint32_t probe_ids(const int (&xs)[K], const int32_t *__restrict__ col,
                                              const CenterSet &C) {
  uint32_t acc = 0;
  if constexpr (kMode == kProbeBitmap) {
    // offsets outside the row's range (and the -2 padding) clamp to the
    // clear sentinel bit: one LDS word read and one bit extract per id;
    // the word's LDS byte address is base + (off >> 5) * 4, one shift and
    // one shift-add (the compiler's own form takes a shift, a mask and an
    // add); v_bfe_u32 takes the bit position from off's low 5 bits
    typedef __attribute__((address_space(3))) const uint32_t lds_u32;
    const uint32_t base =
        (uint32_t)(uintptr_t)(lds_u32 *)reinterpret_cast<const uint32_t *>(C.tbl);
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const uint32_t off = min((uint32_t)(xs[u] - C.lo), C.sentinel);
      uint32_t addr;
      asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(addr) : "v"(off >> 5), "v"(base));
      acc += __builtin_amdgcn_ubfe(*(lds_u32 *)(uintptr_t)addr, off, 1u);
    }
  } else if constexpr (kMode == kProbeHash) {
    // all K home buckets are read before the first compare (one LDS
    // wait), then the few lanes whose home bucket is full without the
    // id continue along the probe sequence; groups of up to four reads
    // keep the register footprint small
    const int4 *__restrict__ tbl = C.tbl;
    constexpr int G = K < 4 ? K : 4;
    uint32_t more = 0;  // bit u: the home bucket of xs[u] is full and lacks it
#pragma unroll
    for (int g = 0; g < K; g += G) {
      int4 vv[G];
#pragma unroll
      for (int j = 0; j < G; ++j) vv[j] = tbl[bucket_of((uint32_t)xs[g + j], C.shift)];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const uint32_t miss = bucket_miss(vv[j], xs[g + j]);
        acc += miss == 0 ? 1u : 0u;
        more |= ((miss != 0) & (vv[j].w != -1)) ? (1u << (g + j)) : 0u;
      }
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < K; ++u)
        if (more & (1u << u))
          acc += table_probe_from(tbl, xs[u], (bucket_of((uint32_t)xs[u], C.shift) + 1) & C.mask,
                                  C.mask);
    }
  } else {
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (xs[u] >= 0) acc += global_probe(col, C.cb, C.dc, xs[u]);
  }
  return acc;
}

// The rows of others[ob:oe] for unroll 1 and 2: one flattened sequence of
// ids, U per lane in flight, each lane following the row boundaries of its
// positions through seg_end / seg_base.
template <class I, int kMode, int U, class O, bool kNarrow>
__device__ __forceinline__ uint64_t stream_others_ids(Rows<I> R, const int32_t *__restrict__ col,
                                                      O others, int64_t ob, int64_t oe,
                                                      int *__restrict__ seg_end,
                                                      int64_t *__restrict__ seg_base,
                                                      const CenterSet &C) {
  const int lane = lane_id();
  uint32_t acc = 0;  // at most chunk * kBigMaxDeg hits per call
  for (int64_t base = ob; base < oe; base += kWave) {
    const int64_t k = base + lane;
    int len = 0;
    int64_t st = 0;
    if (k < oe) others.row(R, k, C.cb, st, len);
    const int inc = wave_inclusive_scan(len);
    const int total = __shfl(inc, kWave - 1);
    seg_end[lane] = inc;
    seg_base[lane] = st - (int64_t)(inc - len);
    wave_lds_sync();
    using Off = IdOffset<kNarrow>;
    int seg = 0;
    int seg_hi = seg_end[0];
    typename Off::T seg_off = (typename Off::T)seg_base[0];
    for (int w0 = 0; w0 < total; w0 += kWave * U) {
      const int t0 = w0 + lane;
      int xs[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * kWave;
        xs[u] = -2;  // past the end
        if (t < total) {
          while (seg_hi <= t) {
            ++seg;
            seg_hi = seg_end[seg];
            seg_off = (typename Off::T)seg_base[seg];
          }
          xs[u] = Off::load(col, seg_off, t);
        }
      }
      acc += probe_ids<kMode, U>(xs, col, C);
    }
    wave_lds_sync();
  }
  return acc;
}

// The rows of others[ob:oe] for unroll 4, 8 and 16.  Each lane takes one
// row of a batch of 64.  A row's 4-id units (the last one partial when the
// row's length is not a multiple of 4) are numbered across the batch, and
// step by step the wavefront takes the next 64 * U / 4 units: lane l loads
// units W + 64 i + l (i < U / 4), each with one 16-byte load, so every load
// instruction reads 1 KiB of consecutive units and a row's last ids are read
// with the rest of their memory line.  A lane finds the row of its unit
// without walking the row ends: at each step every row whose last unit falls
// in the step's window sets its bit in U / 4 LDS words; the row of unit
// W + 64 i + l is the number of rows that ended before W, plus the bits of
// the words before word i, plus the bits of word i below lane l (mbcnt).
// With kTail the ids of a partial unit past its row's end (the next
// row's, read in the same memory line) are replaced by the padding -2, and
// a row whose partial unit would reach past the last id of col[]
// (C.total_ids) reads its last len % 4 ids one by one instead.  Without it
// every row's last len % 4 ids are read one by one by the row's own lane
// before the steps (no masking work; the memory line they share with the
// row's last unit may be fetched twice).
template <class I, int kMode, int U, class O, bool kNarrow, bool kTail>
__device__ __forceinline__ uint64_t stream_others_units(Rows<I> R,
                                                        const int32_t *__restrict__ col,
                                                        O others, int64_t ob, int64_t oe,
                                                        int *__restrict__ seg_end,
                                                        int64_t *__restrict__ seg_base,
                                                        const CenterSet &C) {
  static_assert(U == 4 || U == 8 || U == 16, "unit streaming takes 4, 8 or 16 ids per lane");
  constexpr int IPS = U / 4;        // 16-byte loads per lane per step
  constexpr int WIN = kWave * IPS;  // units per step
  using Off = IdOffset<kNarrow>;
  const int lane = lane_id();
  // the step's row-end bits: 64-bit word i covers units W + 64 i .. + 63
  const uint64_t *ends = reinterpret_cast<const uint64_t *>(seg_end);
  // row ri: seg_base[2 ri] = col position of its unit 0 (unit u at
  // seg_base[2 ri] + 4 u), seg_base[2 ri + 1] = its end in the batch's id
  // numbering (4 * first unit + length)
  uint32_t acc = 0;  // at most chunk * kBigMaxDeg hits per call
  for (int64_t base = ob; base < oe; base += kWave) {
    const int64_t k = base + lane;
    int len = 0;
    int64_t st = 0;
    if (k < oe) others.row(R, k, C.cb, st, len);
    const int nq = (len + 3) >> 2;
    const bool apart = !kTail || st + 4 * (int64_t)nq > C.total_ids;
    const int nu = apart ? (len >> 2) : nq;  // units of this row
    // the last len % 4 ids of a row read apart from its units, one by one
    if (__ballot(apart && (len & 3) != 0) != 0) {
      const int tail = apart ? (len & 3) : 0;
      const typename Off::T p = (typename Off::T)(st + 4 * (int64_t)nu);
      int xs[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) xs[m] = m < tail ? Off::load(col, p, m) : -2;
      acc += probe_ids<kMode, 3>(xs, col, C);
    }
    // rows with units, numbered in lane order
    const uint64_t has = __ballot(nu > 0);
    const int ri = (int)__builtin_amdgcn_mbcnt_hi(
        (uint32_t)(has >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)has, 0u));
    const int inc = wave_inclusive_scan(nu);
    const int T = __shfl(inc, kWave - 1);  // units of the batch
    if (nu > 0) {
      seg_base[2 * ri] = st - 4 * (int64_t)(inc - nu);
      seg_base[2 * ri + 1] = 4 * (int64_t)(inc - nu) + (apart ? 4 * nu : len);
    }
    const int last = inc - 1;  // this row's last unit
    int r0 = 0;                // rows that end before the step's window
    for (int W = 0; W < T; W += WIN) {
      wave_lds_sync();
      if (lane < 2 * IPS) seg_end[lane] = 0;
      wave_lds_sync();
      const int e = last - W;
      if (nu > 0 && (unsigned)e < (unsigned)WIN)
        atomicOr(reinterpret_cast<unsigned *>(&seg_end[e >> 5]), 1u << (e & 31));
      wave_lds_sync();
      int xs[U];
      int lim[IPS];  // ids of each unit inside its row (4 when whole)
      int rs = r0;
#pragma unroll
      for (int i = 0; i < IPS; ++i) {
        const uint64_t ei = ends[i];  // the same word for every lane
        const int u = W + i * kWave + lane;
        int4 v = make_int4(-2, -2, -2, -2);
        lim[i] = 4;
        if (u < T) {
          const int r = rs + (int)__builtin_amdgcn_mbcnt_hi(
                                 (uint32_t)(ei >> 32),
                                 __builtin_amdgcn_mbcnt_lo((uint32_t)ei, 0u));
          const int4 rw = reinterpret_cast<const int4 *>(seg_base)[r];  // base, end
          v = Off::load4(col, (typename Off::T)(((int64_t)rw.y << 32) | (uint32_t)rw.x), 4 * u);
          lim[i] = rw.z - 4 * u;
        }
        rs += __popcll(ei);
        xs[4 * i] = v.x;
        xs[4 * i + 1] = v.y;
        xs[4 * i + 2] = v.z;
        xs[4 * i + 3] = v.w;
      }
      r0 = rs;
      // after every load of the step is issued: the ids past a row's end
      if constexpr (kTail) {
#pragma unroll
        for (int i = 0; i < IPS; ++i) {
          xs[4 * i + 1] = lim[i] > 1 ? xs[4 * i + 1] : -2;
          xs[4 * i + 2] = lim[i] > 2 ? xs[4 * i + 2] : -2;
          xs[4 * i + 3] = lim[i] > 3 ? xs[4 * i + 3] : -2;
        }
      }
      acc += probe_ids<kMode, U>(xs, col, C);
    }
    wave_lds_sync();
  }
  return acc;
}

// Streams the rows of others[ob:oe] and counts the ids found in the
// center's set: in 4-id units for unroll 4, 8 and 16 (the rows' last ids
// with their units with kTail, else apart), id by id below.
template <class I, int kMode, int U, class O, bool kNarrow, bool kTail>
__device__ __forceinline__ uint64_t stream_others(Rows<I> R, const int32_t *__restrict__ col,
                                                  O others, int64_t ob,
                                                  int64_t oe, int *__restrict__ seg_end,
                                                  int64_t *__restrict__ seg_base,
                                                  const CenterSet &C) {
  if constexpr (U >= 4)
    return stream_others_units<I, kMode, U, O, kNarrow, kTail>(R, col, others, ob, oe, seg_end,
                                                                seg_base, C);
  else
    return stream_others_ids<I, kMode, U, O, kNarrow>(R, col, others, ob, oe, seg_end, seg_base,
                                                       C);
}

// stream_others with the id addressing chosen at run time (uniform per
// kernel): 32-bit offsets when the rows lie within 2^30 ids, else 64-bit.
template <class I, int kMode, int U, class O, bool kTail>
__device__ __forceinline__ uint64_t stream_others_rt(bool narrow, Rows<I> R,
                                                     const int32_t *__restrict__ col, O others,
                                                     int64_t ob, int64_t oe,
                                                     int *__restrict__ seg_end,
                                                     int64_t *__restrict__ seg_base,
                                                     const CenterSet &C) {
  return narrow ? stream_others<I, kMode, U, O, true, kTail>(R, col, others, ob, oe, seg_end, seg_base, C)
                : stream_others<I, kMode, U, O, false, kTail>(R, col, others, ob, oe, seg_end, seg_base,
                                                       C);
}

// Builds the LDS set of a center row with the lanes of one wavefront
// (kBlockWide false) or the whole workgroup, choosing the bitmap when the
// row's id range fits the table's bits, else the hash table when the row
// fits it (kMaxHash ids), else the global row; then streams others[ob:oe].
// Every lane (every thread, kBlockWide) calls this with the same item.
template <class I, int U, class O, int kInts, int kMaxHash, bool kBlockWide, bool kTail>
__device__ __forceinline__ uint64_t count_item(bool narrow, Rows<I> R,
                                               const int32_t *__restrict__ col, O others,
                                               int64_t ob, int64_t oe, int *__restrict__ seg_end,
                                               int64_t *__restrict__ seg_base, int *tbl,
                                               int64_t cb, int dc, int64_t total_ids) {
  CenterSet C{reinterpret_cast<const int4 *>(tbl), 0, 0, 0, 0, cb, dc, total_ids};
  if (bitmap_fits<kInts>(col + cb, dc, &C.lo, &C.sentinel)) {
    bitmap_build<kInts, kBlockWide>(tbl, col + cb, dc, C.lo, C.sentinel);
    return stream_others_rt<I, kProbeBitmap, U, O, kTail>(narrow, R, col, others, ob, oe, seg_end,
                                                   seg_base, C);
  }
  if (dc <= kMaxHash) {
    table_build<kInts, kBlockWide>(tbl, col + cb, dc, C.shift, C.mask);
    return stream_others_rt<I, kProbeHash, U, O, kTail>(narrow, R, col, others, ob, oe, seg_end,
                                                 seg_base, C);
  }
  return stream_others_rt<I, kProbeGlobal, U, O, kTail>(narrow, R, col, others, ob, oe, seg_end,
                                                 seg_base, C);
}

// Per-graph counts of a batch: a wavefront keeps per-lane partial sums for
// the graph of its current work item and adds them to that graph's counter
// (one wave reduction, one atomic) only when the graph changes and at the end.
// Consecutive items mostly belong to one graph (graphs own contiguous ids).
struct GraphAcc {
  int g = -1;
  uint64_t acc = 0;
  __device__ __forceinline__ void flush(unsigned long long *__restrict__ count) {
    const uint64_t t = wave_sum(acc);
    if (lane_id() == 0 && t && g >= 0) atomicAdd(&count[g], (unsigned long long)t);
    acc = 0;
  }
  // wave-uniform call: every lane passes the same graph
  __device__ __forceinline__ void add(int graph, uint64_t lane_acc,
                                      unsigned long long *__restrict__ count) {
    if (graph != g) {
      flush(count);
      g = graph;
    }
    acc += lane_acc;
  }
};

// Four wavefronts per SIMD (at most 128 VGPRs): the LDS tables allow four
// workgroups per CU, and the loads in flight need the occupancy.
// item_order: the work items of the wavefront phase, [first, n_items), keyed
// by the col position of their middle other's row (its top 16 bits; any
// coarser order groups the rows as well) above their index j = i - first;
// slots from n_items - first on get a key that sorts last.  *n_keys = the
// number of real keys (the sort reads only those).  The others of an item
// are in row order, so the middle one stands for the rows it streams better
// than the first (a host model of the XCDs' L2s: 14% fewer misses on the
// 117M-edge Chung-Lu graph).
template <class I>
__global__ void k_item_keys(Rows<I> R, OthersPacked others, const int2 *__restrict__ items,
                            const int64_t *__restrict__ cptr, const int64_t *__restrict__ iptr,
                            const int64_t *__restrict__ n_first_runs_p,
                            const int64_t *__restrict__ n_items_p,
                            const unsigned long long *__restrict__ row_stats, int chunk,
                            int64_t max_items, uint64_t *__restrict__ keys,
                            unsigned long long *__restrict__ n_keys) {
  others.resolve();
  const int64_t first = iptr[*n_first_runs_p], n_items = *n_items_p;
  const int sh = max(0, bit_length_u64(row_stats[1]) - 16);
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 == 0) *n_keys = (unsigned long long)max((int64_t)0, n_items - first);
  for (int64_t j = t0; j < max_items; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = first + j;
    uint64_t key = ~0ull;
    if (i < n_items) {
      const int2 it = items[i];
      const int64_t pos = cptr[it.x] + (int64_t)it.y * chunk;
      const int64_t end = min(cptr[it.x + 1], pos + (int64_t)chunk);
      const int64_t mid = pos + (end - pos - 1) / 2;
      const int32_t c = others.center(it.x, pos);
      int64_t st = 0;
      int len = 0;
      others.row(R, mid, R.begin(c), st, len);
      key = ((uint64_t)(st >> sh) << 32) | (uint64_t)j;
    }
    keys[j] = key;
  }
}

// out[i] = items[i] below the wavefront phase, else the item of the i-th
// sorted key.
__global__ void k_item_gather(const int2 *__restrict__ items, const uint64_t *__restrict__ skeys,
                              const int64_t *__restrict__ iptr,
                              const int64_t *__restrict__ n_first_runs_p,
                              const int64_t *__restrict__ n_items_p, int2 *__restrict__ out) {
  const int64_t first = iptr[*n_first_runs_p], n_items = *n_items_p;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_items;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = i < first ? items[i] : items[first + (int64_t)(uint32_t)skeys[i - first]];
}

template <class I, int U, class O, bool kXcdLocal, bool kTail>
__global__ void __launch_bounds__(kBlock, 4)
k_heavy(Rows<I> R, const int32_t *__restrict__ col, const unsigned long long *__restrict__ row_stats,
        const int64_t *__restrict__ cptr, O others, const int2 *__restrict__ items,
        const int64_t *__restrict__ n_items_p, const int64_t *__restrict__ iptr,
        const int64_t *__restrict__ n_huge_runs_p, const int64_t *__restrict__ n_first_runs_p,
        const int64_t *__restrict__ n_big_centers_p, int chunk,
        unsigned long long *__restrict__ next_item, unsigned long long *__restrict__ next_block,
        unsigned long long *__restrict__ xcd_next, int xcd_ticket,
        const int32_t *__restrict__ graph_of, unsigned long long *__restrict__ count) {
  __shared__ int4 tbl_all[kWavesPerBlock][kTblInts / 4];
  __shared__ int seg_end_all[kWavesPerBlock][kWave];
  __shared__ __align__(16) int64_t seg_base_all[kWavesPerBlock][2 * kWave];
  __shared__ int64_t s_ticket;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  others.resolve();
  const bool narrow = rows_narrow(row_stats);
  const int64_t total_ids = (int64_t)row_stats[1];
  int4 *tbl = tbl_all[w];
  int *seg_end = seg_end_all[w];
  int64_t *seg_base = seg_base_all[w];
  const int64_t n_items = *n_items_p;
  // items [0, huge) belong to k_heavy_big; [huge, first) are centers with
  // d+ in [2^(block_min_bits - 1), 4095], taken by whole workgroups over one
  // joined table; then
  // the items of the long centers, one per wavefront ticket; past those a
  // ticket covers kSmallBatch items
  const int64_t huge = iptr[*n_huge_runs_p];
  const int64_t first = iptr[*n_first_runs_p];
  const int64_t n_big = iptr[*n_big_centers_p];
  uint64_t acc_total = 0;
  GraphAcc gacc;

  // Workgroup phase.  Every thread sees the same ticket, so the loop exits
  // together; the barrier at its top keeps the table and the ticket in place
  // until all wavefronts are done with the previous item.  (Batches have no
  // workgroup items.)
  if (first > huge) {
    int *tbl_blk = reinterpret_cast<int *>(&tbl_all[0][0]);
    while (true) {
      __syncthreads();
      if (threadIdx.x == 0) s_ticket = (int64_t)atomicAdd(next_block, 1ull);
      __syncthreads();
      const int64_t bit = huge + s_ticket;
      if (bit >= first) break;
      const int2 item = items[bit];
      const int64_t ob = cptr[item.x] + (int64_t)item.y * chunk;
      const int64_t oe = min(cptr[item.x + 1], ob + (int64_t)chunk);
      const int32_t c = others.center(item.x, ob);
      const int64_t cb = R.begin(c);
      const int dc = (int)(R.end(c) - cb);
      const int64_t part = (oe - ob + kWavesPerBlock - 1) / kWavesPerBlock;
      const int64_t wb = min(oe, ob + w * part), we = min(oe, wb + part);
      const uint64_t acc = count_item<I, U, O, kBlkTblInts, kBlkTblInts / 2 - 1, true, kTail>(
          narrow, R, col, others, wb, we, seg_end, seg_base, tbl_blk, cb, dc, total_ids);
      if (graph_of)
        gacc.add(graph_of[c], acc, count);
      else
        acc_total += acc;
    }
  }

  // Wavefront phase: each ticket gives a range of items [it0, it1).  By
  // default one counter hands out the long centers' items one per ticket,
  // then kSmallBatch per ticket.  In the XCD-local form the items [first,
  // n_items) -- ordered by graph (batches) or by the position of their middle
  // streamed row (item_order) -- are split into kXcds consecutive ranges with
  // one counter each in xcd_next (kXcdStride apart), xcd_ticket items per
  // ticket: workgroup b runs on XCD b % kXcds (round-robin dispatch), so its
  // wavefronts start on that range and the rows its items stream share one
  // L2; a wavefront whose range is used up moves on to the next and leaves
  // after trying all of them.
  [[maybe_unused]] int q = (int)(blockIdx.x % kXcds), tried = 0;
  while (true) {
    int64_t it0, it1;
    unsigned long long ticket = 0;
    if constexpr (kXcdLocal) {
      const int64_t nw = n_items - first;
      const int64_t lo = first + nw * q / kXcds, hi = first + nw * (q + 1) / kXcds;
      if (lane == 0) ticket = atomicAdd(&xcd_next[q * kXcdStride], 1ull);
      it0 = lo + (int64_t)__shfl(ticket, 0) * xcd_ticket;
      it1 = min(it0 + (int64_t)xcd_ticket, hi);
      if (it0 >= hi) {
        q = (q + 1) % kXcds;
        if (++tried == kXcds) break;
        continue;
      }
    } else {
      if (lane == 0) ticket = atomicAdd(next_item, 1ull);
      const int64_t tk = (int64_t)__shfl(ticket, 0);
      if (tk < n_big - first) {
        it0 = first + tk;
        it1 = it0 + 1;
      } else {
        it0 = n_big + (tk - (n_big - first)) * kSmallBatch;
        it1 = min(it0 + (int64_t)kSmallBatch, n_items);
      }
      if (it0 >= n_items) break;
    }
    for (int64_t it = it0; it < it1; ++it) {
      const int2 item = items[it];
      const int64_t ob = cptr[item.x] + (int64_t)item.y * chunk;
      const int64_t oe = min(cptr[item.x + 1], ob + (int64_t)chunk);
      const int32_t c = others.center(item.x, ob);
      const int64_t cb = R.begin(c);
      const int dc = (int)(R.end(c) - cb);
      const uint64_t acc = count_item<I, U, O, kTblInts, kTblMaxDeg, false, kTail>(
          narrow, R, col, others, ob, oe, seg_end, seg_base, reinterpret_cast<int *>(tbl), cb,
          dc, total_ids);
      if (graph_of)
        gacc.add(graph_of[c], acc, count);
      else
        acc_total += acc;
    }
  }
  if (graph_of) {
    gacc.flush(count);
  } else {
    acc_total = wave_sum(acc_total);
    if (lane == 0 && acc_total) atomicAdd(count, (unsigned long long)acc_total);
  }
}

// Per-edge supports (labels for the learned router; per-edge records of the
// reference).  One lane per edge, binary search of the shorter row in the longer.
template <class I>
__global__ void k_edge_supports(Rows<I> R, const int32_t *__restrict__ col,
                                const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                                int64_t E, int32_t *__restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t a0 = R.begin(eu[e]), a1 = R.end(eu[e]), b0 = R.begin(ev[e]), b1 = R.end(ev[e]);
    if (a1 - a0 > b1 - b0) {
      int64_t t0 = a0, t1 = a1;
      a0 = b0; a1 = b1; b0 = t0; b1 = t1;
    }
    int c = 0;
    for (int64_t i = a0; i < a1; ++i) {
      const int x = col[i];
      int64_t lo = b0, hi = b1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < x) lo = mid + 1; else hi = mid;
      }
      if (lo < b1 && col[lo] == x) ++c;
    }
    out[e] = c;
  }
}

// Centers with d+ >= 4096: one workgroup per work item, the center row
// in one block-shared LDS table (up to kBigMaxDeg ids at load <= 1/2), the
// four wavefronts streaming disjoint quarters of the item's others.  Items of
// these centers come first in the heavy-first order; this kernel takes items
// [0, n_first) and k_heavy the rest.
template <class I, int U, class O, bool kTail>
__global__ void __launch_bounds__(kBlock)
k_heavy_big(Rows<I> R, const int32_t *__restrict__ col,
            const unsigned long long *__restrict__ row_stats, const int64_t *__restrict__ cptr,
            O others, const int2 *__restrict__ items, const int64_t *__restrict__ iptr,
            const int64_t *__restrict__ n_first_runs_p, int chunk,
            unsigned long long *__restrict__ next_item, const int32_t *__restrict__ graph_of,
            unsigned long long *__restrict__ count) {
  __shared__ int4 tbl[kBigTblInts / 4];
  __shared__ int seg_end_all[kWavesPerBlock][kWave];
  __shared__ __align__(16) int64_t seg_base_all[kWavesPerBlock][2 * kWave];
  __shared__ int64_t s_item;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  others.resolve();
  const bool narrow = rows_narrow(row_stats);
  const int64_t total_ids = (int64_t)row_stats[1];
  const int64_t n_first = iptr[*n_first_runs_p];
  int *tbl_i = reinterpret_cast<int *>(tbl);
  uint64_t acc_total = 0;
  GraphAcc gacc;
  while (true) {
    if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(next_item, 1ull);
    __syncthreads();
    const int64_t it = s_item;
    __syncthreads();
    if (it >= n_first) break;
    const int2 item = items[it];
    const int64_t ob = cptr[item.x] + (int64_t)item.y * chunk;
    const int64_t oe = min(cptr[item.x + 1], ob + (int64_t)chunk);
    const int32_t c = others.center(item.x, ob);
    const int64_t cb = R.begin(c);
    const int dc = (int)(R.end(c) - cb);
    const int64_t part = (oe - ob + kWavesPerBlock - 1) / kWavesPerBlock;
    const int64_t wb = min(oe, ob + w * part), we = min(oe, wb + part);
    uint64_t acc;
    acc = count_item<I, U, O, kBigTblInts, kBigMaxDeg, true, kTail>(
        narrow, R, col, others, wb, we, seg_end_all[w], seg_base_all[w], tbl_i, cb, dc,
        total_ids);
    __syncthreads();  // the table is rebuilt for the next item
    if (graph_of)
      gacc.add(graph_of[c], acc, count);
    else
      acc_total += acc;
  }
  if (graph_of) {
    gacc.flush(count);
  } else {
    acc_total = wave_sum(acc_total);
    if (lane == 0 && acc_total) atomicAdd(count, (unsigned long long)acc_total);
  }
}

template <class I, class O>
static auto heavy_big_kernel(int unroll, bool tail) -> decltype(&k_heavy_big<I, 4, O, false>) {
  switch (unroll) {
    case 1: return &k_heavy_big<I, 1, O, false>;
    case 2: return &k_heavy_big<I, 2, O, false>;
    case 4: return tail ? &k_heavy_big<I, 4, O, true> : &k_heavy_big<I, 4, O, false>;
    case 8: return tail ? &k_heavy_big<I, 8, O, true> : &k_heavy_big<I, 8, O, false>;
    default: return tail ? &k_heavy_big<I, 16, O, true> : &k_heavy_big<I, 16, O, false>;
  }
}

// k_heavy for an unroll and a form of reading the rows' last ids (tail: in
// masked units; only the unit streaming of unroll 4..16 has the choice)
template <class I, class O, bool kXcdLocal = false>
static auto heavy_kernel(int unroll, bool tail) -> decltype(&k_heavy<I, 4, O, kXcdLocal, false>) {
  switch (unroll) {
    case 1: return &k_heavy<I, 1, O, kXcdLocal, false>;
    case 2: return &k_heavy<I, 2, O, kXcdLocal, false>;
    case 4: return tail ? &k_heavy<I, 4, O, kXcdLocal, true> : &k_heavy<I, 4, O, kXcdLocal, false>;
    case 8: return tail ? &k_heavy<I, 8, O, kXcdLocal, true> : &k_heavy<I, 8, O, kXcdLocal, false>;
    default:
      return tail ? &k_heavy<I, 16, O, kXcdLocal, true> : &k_heavy<I, 16, O, kXcdLocal, false>;
  }
}

__global__ void __launch_bounds__(kBlock)
k_row_stats(const int64_t *__restrict__ rp64, const int32_t *__restrict__ rp32, int64_t n,
            int force_wide, unsigned long long *__restrict__ stats) {
  __shared__ unsigned long long wave_max[kBlock / kWave];
  unsigned long long m = 0;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = rp64 ? rp64[u + 1] - rp64[u] : (int64_t)rp32[u + 1] - rp32[u];
    m = max(m, (unsigned long long)d);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, off, kWave));
  if (lane_id() == 0) wave_max[threadIdx.x / kWave] = m;
  __syncthreads();
  // one same-address atomic per workgroup: thousands per wavefront serialise
  if (threadIdx.x == 0) {
    unsigned long long bm = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) bm = max(bm, wave_max[w]);
    if (bm) atomicMax(&stats[0], bm);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stats[1] = (unsigned long long)(rp64 ? rp64[n] : (int64_t)rp32[n]);
    stats[2] = force_wide ? 1ull : 0ull;
  }
}

// stats[0..2] = (max forward degree, row_ptr[n], force_wide); stats[0] must
// be zero
static int launch_row_stats(const void *row_ptr, int row_ptr_64, int64_t n, bool force_wide,
                            unsigned long long *stats, hipStream_t s) {
  hipLaunchKernelGGL(k_row_stats, dim3(grid_for(n, kBlock, 1024)), dim3(kBlock), 0, s,
                     row_ptr_64 ? (const int64_t *)row_ptr : nullptr,
                     row_ptr_64 ? nullptr : (const int32_t *)row_ptr, n, force_wide ? 1 : 0,
                     stats);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

// ---------------------------------------------------------------------------
// Sorted plan: one packed word per edge -- the center and the other's row (or
// id) for a heavy edge, the sentinel (n << shift) for every other edge -- then
// a radix sort on the center bits whose first pass drops the sentinels groups
// the others of each center (in edge order).  No per-edge global atomics.
// ---------------------------------------------------------------------------
// Learned routing of one wavefront round (kRoute == 2): the round's
// candidate edges are compacted into batch slots (LDS list of their lanes),
// their sampled register statistics read cooperatively (batch_stats), and
// each slot's features and BF16 network output computed in the slot's lane;
// the decision goes back to the candidate's lane.  Every lane of the
// wavefront calls this.
template <int G>
__device__ __forceinline__ bool net_route_round(const RouterNet &net, const float *sp,
                                                const float *snorm, int *s_src, bool cand,
                                                int32_t u, int32_t v, int Du, int Dv) {
  const uint64_t cm = __ballot(cand);
  if (!cm) return false;
  const int lane = lane_id();
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int my_slot = __popcll(cm & lt);
  const int n_c = __popcll(cm);
  if (cand) s_src[my_slot] = lane;
  wave_lds_sync();
  const int src = lane < n_c ? s_src[lane] : 0;
  wave_lds_sync();  // the list is rewritten by the next round
  const int32_t su = __shfl(u, src), sv = __shfl(v, src);
  const int sDu = __shfl(Du, src), sDv = __shfl(Dv, src);
  int32_t a = 0, b = 0;
  if (lane < n_c) {
    a = net.slot ? net.slot[su] : su;
    b = net.slot ? net.slot[sv] : sv;
  }
  const int ms = 1 << net.sample_bits;
  const RegStats st = batch_stats<G, false>(net.regs, 1 << net.p, ms, a, b, n_c);
  int dec = 0;
  if (lane < n_c) {
    const double us = hll_estimate(ms, st.zeros[2], st.z[2]) *
                      (double)(1 << (net.p - net.sample_bits));
    float f[kIn];
    router_features(sDu, sDv, net.card[su], net.card[sv], us, f);
    dec = mlp_prob_bf16(sp, snorm, f) >= net.threshold ? 1 : 0;
  }
  const int back = __shfl(dec, my_slot);
  return cand && back != 0;
}

// kRoute: 0 = plain exact pass, 1 = hybrid routing by the degree rule or a
// route array, 2 = hybrid routing by the learned network (G: lanes per edge of
// its sampled-register read, group_lanes(net.sample_bits)).  kNear: the row
// trimming reads each endpoint's row from L.row_info and searches the cut
// from the edge's own position (count_at_most_near); without it the trimming
// reads L.last_id and searches the whole row (and untrimmed passes keep fewer
// registers).  The kNear instances are held to 64 VGPRs (8 wavefronts per
// SIMD, as the untrimmed pass) so that the search's registers do not cost
// occupancy in this latency-bound pass.
template <class I, int kRoute, int G, bool kNear = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kNear ? 8 : 1)))
k_classify_pack(Rows<I> R, const int32_t *__restrict__ col, const int32_t *__restrict__ eu,
                const int32_t *__restrict__ ev, int64_t E, const unsigned long long *rows_total,
                const uint8_t *__restrict__ skip, int light_max, PackLayout L,
                uint64_t *__restrict__ packed,
                const int32_t *__restrict__ graph_of, unsigned long long *__restrict__ count,
                unsigned long long *__restrict__ n_heavy, int64_t n, HybridRoute hr) {
  uint64_t acc = 0, nh = 0;
  // kRoute: HLL-routed edges, wanted HLL without sketches, network candidates
  uint64_t n_use = 0, n_fb = 0, n_cand = 0;
  const int lane = lane_id();
  const bool rows_layout = L.rows();
  const int lbits = L.lbits();
  const uint64_t sentinel = (uint64_t)n << L.shift;
  const int64_t E_read = edge_limit(E, rows_total);  // edges past it get the sentinel
  float *sp = nullptr, *snorm = nullptr;
  int *s_src = nullptr;
  if constexpr (kRoute == 2) {
    __shared__ float s_params[kParams];
    __shared__ float s_norm[2 * kIn];
    __shared__ int s_list[kBlock / kWave][kWave];
    load_params_bf16(hr.net.params, hr.net.norm, s_params, s_norm);
    sp = s_params;
    snorm = s_norm;
    s_src = s_list[threadIdx.x / kWave];
  }
  const bool trim = rows_layout && (kNear ? L.row_info != nullptr : L.last_id != nullptr);
  const int64_t total_ids = (int64_t)L.stats[1];  // row_ptr[n]: the ids of col
  // the loop bound is uniform per wavefront, so every lane stays active for
  // the cross-lane work below
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < E;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = base + threadIdx.x;
    uint32_t tri = 0;
    int32_t u = 0, v = 0;
    int64_t ub = 0, ue = 0, vb = 0, ve = 0;
    int32_t last_u = 0, last_v = 0;  // with kNear: the rows' largest ids
    const bool active = e < E_read && !(skip && skip[e]);
    if (active) {
      v = ev[e];
      u = eu[e];
      if constexpr (kNear) {
        // launched only with L.row_info set (whatever the packed layout)
        const int4 iu = L.row_info[u], iv = L.row_info[v];
        ub = ((int64_t)iu.y << 32) | (uint32_t)iu.x;
        ue = ub + iu.z;
        last_u = iu.w;
        vb = ((int64_t)iv.y << 32) | (uint32_t)iv.x;
        ve = vb + iv.z;
        last_v = iv.w;
      } else {
        ub = R.begin(u);
        ue = R.end(u);
        vb = R.begin(v);
        ve = R.end(v);
      }
    }
    const int64_t du = ue - ub, dv = ve - vb;
    bool to_hll = false;
    if constexpr (kRoute == 1) {
      if (active) {
        // reads only what decides: D(v) once D(u) reaches gamma, the sketch
        // flags only for edges that want HLL
        bool wants;
        if (hr.route) {
          wants = hr.route[e] != 0;
        } else {
          wants = (hr.degrees ? (int64_t)hr.degrees[u] : du) >= hr.gamma &&
                  (hr.degrees ? (int64_t)hr.degrees[v] : dv) >= hr.gamma;
        }
        const bool both = wants && hr.has[u] == 1 && hr.has[v] == 1;
        to_hll = wants && both;
        n_use += to_hll ? 1 : 0;
        n_fb += (wants && !both) ? 1 : 0;
      }
    }
    if constexpr (kRoute == 2) {
      int Du = 0, Dv = 0;
      bool cand = false;
      if (active && hr.has[u] == 1 && hr.has[v] == 1) {
        Du = hr.degrees ? hr.degrees[u] : (int)du;
        Dv = hr.degrees ? hr.degrees[v] : (int)dv;
        // a sketched vertex without a register row is never a candidate
        cand = (int64_t)min(Du, Dv) >= hr.net.min_cost_degree &&
               (!hr.net.slot || (hr.net.slot[u] >= 0 && hr.net.slot[v] >= 0));
      }
      to_hll = net_route_round<G>(hr.net, sp, snorm, s_src, cand, u, v, Du, Dv);
      n_cand += cand ? 1 : 0;
      n_use += to_hll ? 1 : 0;
    }
    if constexpr (kRoute != 0) {
      if (e < E) hr.use[e] = to_hll ? 1 : 0;
    }
    bool coop = false;           // light edge whose rows fit 32 lanes each
    int la = 0, lb = 0;
    if (e < E) {
      uint64_t word = sentinel;
      if (active && !to_hll && du > 0 && dv > 0) {
        if (du + dv <= light_max) {
          if (du <= kWave / 2 && dv <= kWave / 2) {
            coop = true;
            la = (int)du;
            lb = (int)dv;
          } else {
            tri = merge_count(col, ub, ue, vb, ve);
          }
        } else {
          const bool cv = dv >= du;
          const uint32_t c = (uint32_t)(cv ? v : u), o = (uint32_t)(cv ? u : v);
          int64_t olen = cv ? du : dv;
          // the streamed row is u's own (this pass walks the rows of u, so
          // the search mostly hits cached lines); rows of v are not cut
          if (trim && cv) {
            if constexpr (kNear) {
              if (last_u > last_v) {
                if (total_ids >= kTrimWin) {
                  TrimWindow tw;
                  trim_window_load(col, e, total_ids, tw);
                  olen = count_at_most_near(col, ub, (int)du, last_v, tw, total_ids);
                } else {
                  olen = count_at_most(col, ub, du, last_v);
                }
              }
            } else {
              const int32_t hi = L.last_id[v];
              if (L.last_id[u] > hi) olen = count_at_most(col, ub, du, hi);
            }
          }
          if (olen > 0) {
            word = (uint64_t)c << L.shift;
            if (rows_layout)
              word |= (L.start_field(cv ? ub : vb, cv ? vb : ub) << lbits) | (uint64_t)olen;
            else
              word |= o;
            ++nh;
          }
        }
      }
      packed[e] = word;
    }
    tri += light_intersections(col, coop, ub, la, vb, lb);
    if (graph_of) {
      // one atomic per wavefront when all its triangles belong to one graph
      const int g = tri ? graph_of[u] : -1;
      const uint64_t has = __ballot(g >= 0);
      if (has) {
        const int g0 = __shfl(g, __ffsll((long long)has) - 1);
        if (__ballot(g >= 0 && g != g0) == 0) {
          const uint64_t sum = wave_sum((uint64_t)tri);
          if (lane == 0) atomicAdd(&count[g0], (unsigned long long)sum);
        } else if (g >= 0) {
          atomicAdd(&count[g], (unsigned long long)tri);
        }
      }
    } else {
      acc += tri;
    }
  }
  // one atomic per workgroup and counter: same-address atomics from every
  // wavefront of a large grid serialise
  __shared__ unsigned long long s_red[5][kBlock / kWave];
  nh = wave_sum(nh);
  acc = wave_sum(acc);
  if constexpr (kRoute != 0) {
    n_use = wave_sum(n_use);
    n_fb = wave_sum(n_fb);
    n_cand = wave_sum(n_cand);
  }
  if (lane == 0) {
    s_red[0][threadIdx.x / kWave] = nh;
    s_red[1][threadIdx.x / kWave] = acc;
    s_red[2][threadIdx.x / kWave] = n_use;
    s_red[3][threadIdx.x / kWave] = n_fb;
    s_red[4][threadIdx.x / kWave] = n_cand;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t[5] = {0, 0, 0, 0, 0};
    for (int w = 0; w < kBlock / kWave; ++w)
      for (int q = 0; q < 5; ++q) t[q] += s_red[q][w];
    if (t[0]) atomicAdd(n_heavy, t[0]);
    if (!graph_of && t[1]) atomicAdd(count, t[1]);
    if constexpr (kRoute != 0) {
      if (t[2]) atomicAdd(&hr.counters[0], t[2]);
      if (t[3]) atomicAdd(&hr.counters[1], t[3]);
      if (kRoute == 2 && t[4]) atomicAdd(&hr.counters[3], t[4]);
    }
  }
}

// run offsets (int64, R + 1 entries) and the bucket key of each run's center
template <class I>
__global__ void k_run_offsets(Rows<I> R, const uint64_t *__restrict__ s,
                              const int32_t *__restrict__ run_idx, const int64_t *__restrict__ R_p,
                              const unsigned long long *__restrict__ n_heavy_p, int64_t max_runs,
                              int shift, int64_t *__restrict__ run_ptr, uint8_t *__restrict__ key) {
  const int64_t nr = *R_p;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= max_runs;
       r += (int64_t)gridDim.x * blockDim.x) {
    if (r < nr) {
      const int64_t b = run_idx[r];
      run_ptr[r] = b;
      const uint64_t d = (uint64_t)R.deg((int64_t)(s[b] >> shift));
      key[r] = (uint8_t)(kCenterBuckets - (64 - __clzll((long long)d)));
    } else {
      if (r == nr) run_ptr[r] = (int64_t)*n_heavy_p;
      if (r < max_runs) key[r] = 255;
    }
  }
}

// Sort keys that order runs by graph: (graph of the run's center << 32 | run)
// for r < n_runs.
__global__ void k_run_graph_keys(const uint64_t *__restrict__ s,
                                 const int64_t *__restrict__ run_ptr,
                                 const int64_t *__restrict__ n_runs_p, int shift,
                                 const int32_t *__restrict__ graph_of,
                                 uint64_t *__restrict__ keys) {
  const int64_t nr = *n_runs_p;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nr;
       r += (int64_t)gridDim.x * blockDim.x)
    keys[r] = ((uint64_t)(uint32_t)graph_of[s[run_ptr[r]] >> shift] << 32) | (uint64_t)r;
}

// perm[i] = run of the i-th sorted key (i < n_runs); the bucket boundaries
// put every run in the wavefront phase (starts[0 .. kCenterBuckets - 1] = 0,
// starts[kCenterBuckets] = n_runs).
__global__ void k_graph_perm(const uint64_t *__restrict__ sorted_keys,
                             const int64_t *__restrict__ n_runs_p, int32_t *__restrict__ perm,
                             int64_t *__restrict__ starts) {
  const int64_t nr = *n_runs_p;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < nr; i += (int64_t)gridDim.x * blockDim.x)
    perm[i] = (int32_t)(uint32_t)sorted_keys[i];
  if (t0 <= kCenterBuckets) starts[t0] = t0 == kCenterBuckets ? nr : 0;
}

__global__ void k_run_item_counts(const int64_t *__restrict__ run_ptr,
                                  const int32_t *__restrict__ perm,
                                  const int64_t *__restrict__ n_runs_p, int64_t max_runs, int chunk,
                                  int32_t *__restrict__ nitems) {
  // entries from n_runs on are not read (the scan is bounded by n_runs)
  const int64_t nr = min(*n_runs_p, max_runs);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = perm[i];
    nitems[i] = (int)((run_ptr[r + 1] - run_ptr[r] + chunk - 1) / chunk);
  }
}

// Others per work item: at most 65536 keeps each lane's 32-bit hit count exact.
static int effective_chunk(const tc_exact_params *params) {
  return (params && params->chunk > 0) ? std::min(params->chunk, 65536) : 512;
}

// Scratch bytes exact_sorted_plan carves for n vertices and E edges: its own
// pieces in allocation order, then those of the radix sort, the run
// compaction, the bucket sort and the item scan (each rounded by arena_round).
static size_t exact_sorted_scratch_bytes(int64_t n, int64_t E, int chunk, bool csr_mode,
                                         bool item_order) {
  int key_bits = 1;
  while ((int64_t(1) << key_bits) <= n) ++key_bits;
  const int64_t max_items = n + E / chunk + 1;
  const int64_t pieces[] = {
      csr_mode ? (int64_t)sizeof(int32_t) * E : 0,          // eu_buf
      (int64_t)sizeof(uint64_t) * E,                         // packed
      (int64_t)sizeof(uint64_t) * E,                         // tmp
      E,                                                     // flags
      (int64_t)sizeof(int32_t) * n,                          // run_idx
      (int64_t)sizeof(int64_t),                              // n_runs
      (int64_t)sizeof(int64_t) * (n + 1),                    // run_ptr
      n,                                                     // rkey
      (int64_t)sizeof(int32_t) * n,                          // perm
      (int64_t)sizeof(int64_t) * (kCenterBuckets + 1),       // starts
      (int64_t)sizeof(int32_t) * n,                          // nitems
      (int64_t)sizeof(int64_t) * (n + 1),                    // iptr
      (int64_t)sizeof(int2) * max_items,                     // items
      (int64_t)(7 * sizeof(unsigned long long)),             // ctr
      item_order ? (int64_t)sizeof(uint64_t) * max_items : 0,      // ikeys
      item_order ? (int64_t)sizeof(uint64_t) * max_items : 0,      // ikeys2
      item_order ? (int64_t)sizeof(int2) * max_items : 0,          // items2
      item_order ? (int64_t)sizeof(unsigned long long) : 0,        // n_keys
      item_order ? (int64_t)sizeof(unsigned long long) * kXcds * kXcdStride : 0,  // xcd_ctr
  };
  size_t b = 0;
  for (const int64_t x : pieces) b += x > 0 ? arena_round((size_t)x) : 0;
  b += radix_sort_scratch_bytes(E, 64 - key_bits, 64, true);
  if (item_order) b += radix_sort_scratch_bytes(max_items, 32, 48);
  b += compact_scratch_bytes(E);
  b += bucket_sort_scratch_bytes(n, kCenterBuckets);
  b += scan_scratch_bytes(n);
  return b;
}

static int exact_sorted_plan(const void *row_ptr, int row_ptr_64, const int32_t *col, int64_t n,
                             const int32_t *eu, const int32_t *ev, int64_t E,
                             const uint8_t *skip, const int32_t *graph_of,
                             unsigned long long *count, int light_max, int chunk, int wpc,
                             int unroll, int blk_bits, bool pack_ids, bool wide, bool trim,
                             bool tail_units, bool item_order, const BatchPlan &bp,
                             hipStream_t s, const HybridRoute *hr) {
  Scratch eu_buf;
  const bool csr_mode = eu == nullptr;
  if (!eu) {
    TC_TRY(eu_buf.alloc(sizeof(int32_t) * E, s));
    TC_DISPATCH_ROWS(row_ptr_64, I, {
      hipLaunchKernelGGL(k_expand_rows<I>, dim3(grid_for(n * kWave, kBlock, 65536)),
                         dim3(kBlock), 0, s, (const I *)row_ptr, n, E, eu_buf.as<int32_t>());
    });
    TC_HIP(hipGetLastError());
    eu = eu_buf.as<int32_t>();
    ev = col;
  }
  int key_bits = 1;
  while ((int64_t(1) << key_bits) <= n) ++key_bits;  // center ids and the sentinel n
  // packed words: the center in the top key_bits bits (PackLayout)
  const int shift = 64 - key_bits;
  const int64_t max_runs = n;                         // at most one run per center
  Scratch packed, tmp, flags, run_idx, n_runs, run_ptr, rkey, perm, starts, nitems, iptr,
      items, ctr;
  TC_TRY(packed.alloc(sizeof(uint64_t) * E, s));
  TC_TRY(tmp.alloc(sizeof(uint64_t) * E, s));
  TC_TRY(flags.alloc(E, s));
  TC_TRY(run_idx.alloc(sizeof(int32_t) * max_runs, s));
  TC_TRY(n_runs.alloc(sizeof(int64_t), s));
  TC_TRY(run_ptr.alloc(sizeof(int64_t) * (max_runs + 1), s));
  TC_TRY(rkey.alloc(max_runs, s));
  TC_TRY(perm.alloc(sizeof(int32_t) * max_runs, s));
  TC_TRY(starts.alloc(sizeof(int64_t) * (kCenterBuckets + 1), s));
  TC_TRY(nitems.alloc(sizeof(int32_t) * max_runs, s));
  TC_TRY(iptr.alloc(sizeof(int64_t) * (max_runs + 1), s));
  const int64_t max_items = max_runs + E / chunk + 1;
  TC_TRY(items.alloc(sizeof(int2) * max_items, s));
  TC_TRY(ctr.alloc(7 * sizeof(unsigned long long), s));
  // batches: runs ordered by graph, one item queue per XCD
  const bool xcd_local = bp.graph_bits > 0 && graph_of != nullptr;
  // item_order (one graph): the wavefront phase's items sorted by the
  // position of their middle streamed row, handed out from one queue per XCD
  const bool band = item_order && !xcd_local && graph_of == nullptr;
  Scratch xcd_ctr, ikeys, ikeys2, items2, n_keys;
  if (band) {
    TC_TRY(ikeys.alloc(sizeof(uint64_t) * max_items, s));
    TC_TRY(ikeys2.alloc(sizeof(uint64_t) * max_items, s));
    TC_TRY(items2.alloc(sizeof(int2) * max_items, s));
    TC_TRY(n_keys.alloc(sizeof(unsigned long long), s));
  }
  if (xcd_local || band) {
    TC_TRY(xcd_ctr.alloc(sizeof(unsigned long long) * kXcds * kXcdStride, s));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kBlock), 0, s, xcd_ctr.as<unsigned long long>(),
                       (int64_t)kXcds * kXcdStride, 0ull);
  }
  unsigned long long *next_item = ctr.as<unsigned long long>();
  unsigned long long *n_heavy = next_item + 1;
  unsigned long long *row_stats = next_item + 4;  // 3 entries
  // with trim_rows each vertex's row start, length and last id go to tmp as
  // one 16-byte entry (tmp is the radix sort's second buffer, unused until
  // the sort after the edge pass; it holds them when E >= 2n); the learned
  // router's edge pass, and graphs with fewer edges, take the rows' last ids
  // alone from nitems (int32[n], unused until the planner after the sort)
  const bool trim_on = trim && !pack_ids;
  const bool router = hr && hr->net.params;
  const bool near = trim_on && !router &&
                    (int64_t)sizeof(uint64_t) * E >= (int64_t)sizeof(int4) * n;
  PackLayout layout{shift, pack_ids ? 1 : 0, row_stats, bp.rel_bits,
                    trim_on && !near ? nitems.as<int32_t>() : nullptr};
  layout.row_info = near ? tmp.as<int4>() : nullptr;
  const unsigned long long *rows_total = csr_mode ? row_stats + 1 : nullptr;

  const unsigned gr = grid_for(max_runs + 1, kBlock, 16384);
  const unsigned ge = grid_for(E, kBlock, 16384);
  unsigned long long *next_big = next_item + 2;
  unsigned long long *next_blk = next_item + 3;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, next_item, 7, 0ull);
  TC_TRY(launch_row_stats(row_ptr, row_ptr_64, n, wide, row_stats, s));
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    Rows<I> R{(const I *)row_ptr};
    {
      StageTimer _timer(TC_STAGE_CLASSIFY, s);
      if (layout.row_info)
        hipLaunchKernelGGL(k_row_info<I>, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0, s,
                           R, col, n, tmp.as<int4>());
      if (layout.last_id)
        hipLaunchKernelGGL(k_last_ids<I>, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0, s,
                           R, col, n, nitems.as<int32_t>());
      if (hr && hr->net.params) {
        TC_DISPATCH_GROUP(hr->net.sample_bits, G, {
          hipLaunchKernelGGL((k_classify_pack<I, 2, G>), dim3(ge), dim3(kBlock), 0, s, R, col,
                             eu, ev, E, rows_total, skip, light_max, layout,
                             packed.as<uint64_t>(), graph_of, count, n_heavy, n, *hr);
        });
      } else if (hr) {
        const auto kfn = layout.row_info ? k_classify_pack<I, 1, 1, true>
                                        : k_classify_pack<I, 1, 1, false>;
        hipLaunchKernelGGL(kfn, dim3(ge), dim3(kBlock), 0, s, R, col, eu, ev, E, rows_total, skip,
                           light_max, layout, packed.as<uint64_t>(), graph_of, count, n_heavy, n,
                           *hr);
      } else {
        const auto kfn = layout.row_info ? k_classify_pack<I, 0, 1, true>
                                        : k_classify_pack<I, 0, 1, false>;
        hipLaunchKernelGGL(kfn, dim3(ge), dim3(kBlock), 0, s, R, col, eu, ev, E, rows_total, skip,
                           light_max, layout, packed.as<uint64_t>(), graph_of, count, n_heavy, n,
                           HybridRoute{});
      }
    }
    TC_HIP(hipGetLastError());
    // Small inputs are launch-bound: when classify found no heavy edge, the
    // sort, planner and heavy kernels (about 30 launches) have nothing to do.
    // Outside stream capture and graph building, read the heavy count back
    // once and stop; large inputs never wait on the host.
    if (E <= kEarlyExitMaxEdges && current_arena() == nullptr) {
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      TC_HIP(hipStreamIsCapturing(s, &cap));
      if (cap == hipStreamCaptureStatusNone) {
        unsigned long long h_heavy = 0;
        TC_HIP(hipMemcpyAsync(&h_heavy, n_heavy, sizeof(h_heavy), hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        if (h_heavy == 0) return TC_OK;
      }
    }
    uint64_t *sorted = nullptr;
    // the first pass leaves out the light edges' sentinel words; the later
    // passes sort the n_heavy heavy words only; the last pass flags the first
    // word of each center's run
    const uint64_t sentinel = (uint64_t)n << shift;
    // With sort_bits, only the center's low sort_bits bits are sorted: the
    // caller guarantees that words of centers with equal low bits never
    // interleave in packed order (a batch whose edges are grouped by graph,
    // each graph at most 2^sort_bits vertices), so each center's words still
    // form one run.  The run flags compare every center bit, so a run never
    // mixes centers; the compaction keeps at most max_runs runs whatever the
    // input, so the run arrays cannot overflow.
    const int sort_hi = (bp.sort_bits > 0 && bp.sort_bits < key_bits) ? shift + bp.sort_bits
                                                                       : 64;
    TC_TRY(radix_sort_u64(packed.as<uint64_t>(), tmp.as<uint64_t>(), E, shift, sort_hi, &sorted,
                          s, &sentinel, n_heavy, flags.as<uint8_t>()));
    TC_TRY(compact_indices(flags.as<uint8_t>(), E, run_idx.as<int32_t>(), n_runs.as<int64_t>(),
                           s, max_runs));
    hipLaunchKernelGGL(k_run_offsets<I>, dim3(gr), dim3(kBlock), 0, s, R,
                       (const uint64_t *)sorted, (const int32_t *)run_idx.as<int32_t>(),
                       (const int64_t *)n_runs.as<int64_t>(), (const unsigned long long *)n_heavy,
                       max_runs, shift, run_ptr.as<int64_t>(), rkey.as<uint8_t>());
    TC_HIP(hipGetLastError());
    if (xcd_local) {
      // runs ordered by graph (stable: by center inside a graph)
      Scratch gk, gk2;
      TC_TRY(gk.alloc(sizeof(uint64_t) * max_runs, s));
      TC_TRY(gk2.alloc(sizeof(uint64_t) * max_runs, s));
      hipLaunchKernelGGL(k_run_graph_keys, dim3(gr), dim3(kBlock), 0, s, (const uint64_t *)sorted,
                         (const int64_t *)run_ptr.as<int64_t>(),
                         (const int64_t *)n_runs.as<int64_t>(), shift, graph_of,
                         gk.as<uint64_t>());
      TC_HIP(hipGetLastError());
      uint64_t *gsorted = nullptr;
      TC_TRY(radix_sort_u64(gk.as<uint64_t>(), gk2.as<uint64_t>(), max_runs, 32,
                            32 + bp.graph_bits, &gsorted, s, nullptr,
                            (const unsigned long long *)n_runs.as<int64_t>()));
      hipLaunchKernelGGL(k_graph_perm, dim3(gr), dim3(kBlock), 0, s, (const uint64_t *)gsorted,
                         (const int64_t *)n_runs.as<int64_t>(), perm.as<int32_t>(),
                         starts.as<int64_t>());
      TC_HIP(hipGetLastError());
    } else {
      // runs ordered longest center row first
      TC_TRY(bucket_sort_u8(rkey.as<uint8_t>(), max_runs, kCenterBuckets, perm.as<int32_t>(),
                            starts.as<int64_t>(), s));
    }
    const int64_t *n_sorted_runs = starts.as<int64_t>() + kCenterBuckets;
    const int64_t *n_big_runs = starts.as<int64_t>() + (kCenterBuckets - kBigCenterBits + 1);
    hipLaunchKernelGGL(k_run_item_counts, dim3(gr), dim3(kBlock), 0, s,
                       (const int64_t *)run_ptr.as<int64_t>(), (const int32_t *)perm.as<int32_t>(),
                       n_sorted_runs, max_runs, chunk, nitems.as<int32_t>());
    TC_TRY(exclusive_scan<int32_t>(nitems.as<int32_t>(), iptr.as<int64_t>(), max_runs, nullptr,
                                   s, n_sorted_runs));
    hipLaunchKernelGGL(k_item_fill, dim3(gr), dim3(kBlock), 0, s,
                       (const int32_t *)nitems.as<int32_t>(), (const int32_t *)perm.as<int32_t>(),
                       (const int64_t *)iptr.as<int64_t>(), n_sorted_runs, items.as<int2>());
    TC_HIP(hipGetLastError());
    const unsigned gh = (unsigned)std::max(1, num_cus() * wpc / kWavesPerBlock);
    const int64_t *n_huge = starts.as<int64_t>() + (kCenterBuckets - kBigMinBits + 1);
    const int64_t *n_first = starts.as<int64_t>() + (kCenterBuckets - blk_bits + 1);
    const int64_t *n_items_d = iptr.as<int64_t>() + max_runs;
    const int2 *heavy_items = items.as<int2>();
    if (band) {
      const unsigned gi = grid_for(max_items, kBlock, 16384);
      hipLaunchKernelGGL(k_item_keys<I>, dim3(gi), dim3(kBlock), 0, s, R,
                         OthersPacked{(const uint64_t *)sorted, layout},
                         (const int2 *)items.as<int2>(), (const int64_t *)run_ptr.as<int64_t>(),
                         (const int64_t *)iptr.as<int64_t>(), n_first, n_items_d,
                         (const unsigned long long *)row_stats, chunk, max_items,
                         ikeys.as<uint64_t>(), n_keys.as<unsigned long long>());
      TC_HIP(hipGetLastError());
      uint64_t *skeys = nullptr;
      TC_TRY(radix_sort_u64(ikeys.as<uint64_t>(), ikeys2.as<uint64_t>(), max_items, 32, 48,
                            &skeys, s, nullptr, n_keys.as<unsigned long long>()));
      hipLaunchKernelGGL(k_item_gather, dim3(gi), dim3(kBlock), 0, s,
                         (const int2 *)items.as<int2>(), (const uint64_t *)skeys,
                         (const int64_t *)iptr.as<int64_t>(), n_first, n_items_d,
                         items2.as<int2>());
      TC_HIP(hipGetLastError());
      heavy_items = items2.as<int2>();
    }
    {
      StageTimer _timer(TC_STAGE_HEAVY, s);
      const auto kbig = heavy_big_kernel<I, OthersPacked>(unroll, tail_units);
      hipLaunchKernelGGL(kbig, dim3(num_cus()), dim3(kBlock), 0, s, R, col,
                         (const unsigned long long *)row_stats,
                         (const int64_t *)run_ptr.as<int64_t>(),
                         OthersPacked{(const uint64_t *)sorted, layout},
                         (const int2 *)items.as<int2>(),
                         (const int64_t *)iptr.as<int64_t>(), n_huge, chunk, next_big, graph_of,
                         count);
      const auto kfn = (xcd_local || band)
                           ? heavy_kernel<I, OthersPacked, true>(unroll, tail_units)
                           : heavy_kernel<I, OthersPacked>(unroll, tail_units);
      hipLaunchKernelGGL(kfn, dim3(gh), dim3(kBlock), 0, s, R,
                         col, (const unsigned long long *)row_stats,
                         (const int64_t *)run_ptr.as<int64_t>(),
                         OthersPacked{(const uint64_t *)sorted, layout}, heavy_items, n_items_d,
                         (const int64_t *)iptr.as<int64_t>(), n_huge, n_first, n_big_runs, chunk,
                         next_item, next_blk,
                         (xcd_local || band) ? xcd_ctr.as<unsigned long long>() : nullptr,
                         band ? kBandTicket : kBatchTicket, graph_of, count);
    }
    TC_HIP(hipGetLastError());
  });
  return TC_OK;
}

int exact_count_impl(const void *row_ptr, int row_ptr_64, const int32_t *col, int64_t n,
                     const int32_t *eu, const int32_t *ev, int64_t E, const uint8_t *skip,
                     const int32_t *graph_of, unsigned long long *count,
                     const tc_exact_params *params, hipStream_t s, const HybridRoute *hr,
                     const BatchPlan *bp) {
  const int light_max = (params && params->light_max > 0) ? params->light_max : 32;
  const int chunk = effective_chunk(params);
  const int wpc = (params && params->waves_per_cu > 0) ? params->waves_per_cu : 16;
  const int unroll = (params && params->unroll > 0) ? params->unroll : kDefaultUnroll;
  TC_CHECK(unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8 || unroll == 16,
           TC_ERR_INVALID, "unroll must be 1, 2, 4, 8 or 16, got %d", unroll);
  const int plan = params ? params->plan : TC_PLAN_SORTED;
  const int blk_bits = (params && params->block_min_bits > 0) ? params->block_min_bits
                                                              : kBlkMinBits;
  TC_CHECK(blk_bits >= kBigCenterBits && blk_bits <= kBigMinBits, TC_ERR_INVALID,
           "block_min_bits must be in [%d, %d], got %d", kBigCenterBits, kBigMinBits, blk_bits);
  TC_CHECK(!hr || (plan == TC_PLAN_SORTED && !skip && eu && ev), TC_ERR_INVALID,
           "hybrid routing in the exact pass needs the sorted plan, an edge list and no skip mask");
  const bool tail_units = params && params->tail_units != 0;
  if (E <= 0 || n <= 0) return TC_OK;
  if (plan == TC_PLAN_SORTED)
    return exact_sorted_plan(row_ptr, row_ptr_64, col, n, eu, ev, E, skip, graph_of, count,
                             light_max, chunk, wpc, unroll, blk_bits,
                             params && params->pack_ids != 0,
                             params && params->wide_offsets != 0,
                             params && params->trim_rows != 0, tail_units,
                             params && params->item_order != 0, bp ? *bp : BatchPlan{}, s, hr);

  Scratch eu_buf;
  const bool csr_mode = eu == nullptr;
  if (!eu) {
    TC_TRY(eu_buf.alloc(sizeof(int32_t) * E, s));
    TC_DISPATCH_ROWS(row_ptr_64, I, {
      hipLaunchKernelGGL(k_expand_rows<I>, dim3(grid_for(n * kWave, kBlock, 65536)),
                         dim3(kBlock), 0, s, (const I *)row_ptr, n, E, eu_buf.as<int32_t>());
    });
    TC_HIP(hipGetLastError());
    eu = eu_buf.as<int32_t>();
    ev = col;
  }

  Scratch cnt, cptr, others, nitems, iptr, items, ctr, ckey, perm, starts;
  TC_TRY(ckey.alloc(n, s));
  TC_TRY(perm.alloc(sizeof(int32_t) * n, s));
  TC_TRY(starts.alloc(sizeof(int64_t) * (kCenterBuckets + 1), s));
  TC_TRY(cnt.alloc(sizeof(int32_t) * n, s));
  TC_TRY(cptr.alloc(sizeof(int64_t) * (n + 1), s));
  TC_TRY(others.alloc(sizeof(int32_t) * E, s));
  TC_TRY(nitems.alloc(sizeof(int32_t) * n, s));
  TC_TRY(iptr.alloc(sizeof(int64_t) * (n + 1), s));
  const int64_t max_items = n + E / chunk + 1;
  TC_TRY(items.alloc(sizeof(int2) * max_items, s));
  TC_TRY(ctr.alloc(6 * sizeof(unsigned long long), s));
  unsigned long long *row_stats = ctr.as<unsigned long long>() + 3;  // 3 entries
  const unsigned long long *rows_total = csr_mode ? row_stats + 1 : nullptr;

  const unsigned gn = grid_for(n, kBlock, 16384);
  const unsigned ge = grid_for(E, kBlock, 16384);
  hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(kBlock), 0, s, cnt.as<int32_t>(), n, 0);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, ctr.as<unsigned long long>(), 6, 0ull);
  TC_TRY(launch_row_stats(row_ptr, row_ptr_64, n, params && params->wide_offsets != 0,
                          row_stats, s));
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    Rows<I> R{(const I *)row_ptr};
    {
      StageTimer _timer(TC_STAGE_CLASSIFY, s);
      hipLaunchKernelGGL(k_classify<I>, dim3(ge), dim3(kBlock), 0, s, R, col, eu, ev, E,
                         rows_total, skip, light_max, cnt.as<int32_t>(), graph_of, count);
    }
    TC_HIP(hipGetLastError());
    TC_TRY(exclusive_scan<int32_t>(cnt.as<int32_t>(), cptr.as<int64_t>(), n, nullptr, s));
    // reuse nitems as the scatter cursor, then as the per-center item count
    hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(kBlock), 0, s, nitems.as<int32_t>(), n, 0);
    hipLaunchKernelGGL(k_scatter_others<I>, dim3(ge), dim3(kBlock), 0, s, R, eu, ev, E,
                       rows_total, skip, light_max, (const int64_t *)cptr.as<int64_t>(), nitems.as<int32_t>(),
                       others.as<int32_t>());
    TC_HIP(hipGetLastError());
    // order centers longest row first (stable bucket sort on bit length)
    hipLaunchKernelGGL(k_center_keys<I>, dim3(gn), dim3(kBlock), 0, s, R,
                       (const int32_t *)cnt.as<int32_t>(), n, ckey.as<uint8_t>());
    TC_HIP(hipGetLastError());
    TC_TRY(bucket_sort_u8(ckey.as<uint8_t>(), n, kCenterBuckets, perm.as<int32_t>(),
                          starts.as<int64_t>(), s));
    const int64_t *n_cent = starts.as<int64_t>() + kCenterBuckets;
    const int64_t *n_big_cent = starts.as<int64_t>() + (kCenterBuckets - kBigCenterBits + 1);
    hipLaunchKernelGGL(k_item_counts, dim3(gn), dim3(kBlock), 0, s,
                       (const int32_t *)cnt.as<int32_t>(), (const int32_t *)perm.as<int32_t>(),
                       n_cent, n, chunk, nitems.as<int32_t>());
    TC_TRY(exclusive_scan<int32_t>(nitems.as<int32_t>(), iptr.as<int64_t>(), n, nullptr, s));
    hipLaunchKernelGGL(k_item_fill, dim3(gn), dim3(kBlock), 0, s,
                       (const int32_t *)nitems.as<int32_t>(), (const int32_t *)perm.as<int32_t>(),
                       (const int64_t *)iptr.as<int64_t>(), n_cent, items.as<int2>());
    TC_HIP(hipGetLastError());
    const unsigned gh = (unsigned)std::max(1, num_cus() * wpc / kWavesPerBlock);
    {
      StageTimer _timer(TC_STAGE_HEAVY, s);
      const int64_t *n_huge = starts.as<int64_t>() + (kCenterBuckets - kBigMinBits + 1);
      const int64_t *n_first = starts.as<int64_t>() + (kCenterBuckets - blk_bits + 1);
      const auto kbig = heavy_big_kernel<I, OthersIds>(unroll, tail_units);
      hipLaunchKernelGGL(kbig, dim3(num_cus()), dim3(kBlock), 0, s, R, col,
                         (const unsigned long long *)row_stats,
                         (const int64_t *)cptr.as<int64_t>(),
                         OthersIds{(const int32_t *)others.as<int32_t>()},
                         (const int2 *)items.as<int2>(), (const int64_t *)iptr.as<int64_t>(),
                         n_huge, chunk, ctr.as<unsigned long long>() + 1, graph_of, count);
      const auto kfn = heavy_kernel<I, OthersIds>(unroll, tail_units);
      hipLaunchKernelGGL(kfn, dim3(gh), dim3(kBlock), 0, s, R, col,
                         (const unsigned long long *)row_stats,
                         (const int64_t *)cptr.as<int64_t>(),
                         OthersIds{(const int32_t *)others.as<int32_t>()},
                         (const int2 *)items.as<int2>(), (const int64_t *)(iptr.as<int64_t>() + n),
                         (const int64_t *)iptr.as<int64_t>(), n_huge, n_first, n_big_cent, chunk,
                         ctr.as<unsigned long long>(), ctr.as<unsigned long long>() + 2, nullptr,
                         kSmallBatch /* unused: no XCD queues */, graph_of, count);
    }
    TC_HIP(hipGetLastError());
  });
  return TC_OK;
}

}  // namespace tc

extern "C" int tc_exact_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                              int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                              int64_t n_edges, uint64_t *count, const tc_exact_params *params,
                              void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(row_ptr && count, TC_ERR_INVALID, "row_ptr and count are required");
  TC_CHECK(n_nodes >= 0 && n_edges >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK((edge_u == nullptr) == (edge_v == nullptr), TC_ERR_INVALID,
           "edge_u and edge_v must both be given or both be NULL");
  TC_CHECK(n_edges == 0 || col_idx, TC_ERR_INVALID, "col_idx is required");
  TC_CHECK(n_nodes < INT32_MAX, TC_ERR_INVALID, "node ids must fit int32");
  TC_CHECK(n_edges < INT32_MAX, TC_ERR_INVALID, "edge positions must fit int32");
  TC_TRY(ensure_init());
  return exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges,
                          nullptr, nullptr, (unsigned long long *)count, params,
                          (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// The exact count without allocation: the pipeline's scratch is carved from
// one file-scope device array (per device, held until the process ends)
// instead of the stream workspace or the memory pool.
// ---------------------------------------------------------------------------
namespace tc {

constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;  // 128 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static char *g_scratch_ptr[64] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device (the array's name is not a device address in host code).
static int scratch_base(char **out) {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  TC_CHECK(dev >= 0 && dev < 64, TC_ERR_UNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TC_CHECK(p != nullptr, TC_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return TC_OK;
}

}  // namespace tc

extern "C" int tc_exact_count_noalloc(const void *row_ptr, int row_ptr_64,
                                      const int32_t *col_idx, int64_t n_nodes,
                                      const int32_t *edge_u, const int32_t *edge_v,
                                      int64_t n_edges, uint64_t *count,
                                      const tc_exact_params *params, void *stream) {
  using namespace tc;
  TC_CHECK(row_ptr && count, TC_ERR_INVALID, "row_ptr and count are required");
  TC_CHECK(n_nodes >= 0 && n_edges >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK((edge_u == nullptr) == (edge_v == nullptr), TC_ERR_INVALID,
           "edge_u and edge_v must both be given or both be NULL");
  TC_CHECK(n_edges == 0 || col_idx, TC_ERR_INVALID, "col_idx is required");
  TC_CHECK(n_nodes < INT32_MAX, TC_ERR_INVALID, "node ids must fit int32");
  TC_CHECK(n_edges < INT32_MAX, TC_ERR_INVALID, "edge positions must fit int32");
  TC_CHECK(!params || params->plan == TC_PLAN_SORTED, TC_ERR_INVALID,
           "tc_exact_count_noalloc runs the sorted plan only");
  TC_TRY(ensure_init());
  if (n_edges == 0 || n_nodes == 0) return TC_OK;
  const size_t need =
      exact_sorted_scratch_bytes(n_nodes, n_edges, effective_chunk(params), edge_u == nullptr,
                                 params && params->item_order != 0);
  TC_CHECK(need <= kNoAllocScratchBytes, TC_ERR_INVALID,
           "%lld nodes and %lld edges need %zu bytes of scratch; tc_exact_count_noalloc holds "
           "%zu (use tc_exact_count for this graph)",
           (long long)n_nodes, (long long)n_edges, need, kNoAllocScratchBytes);
  char *base = nullptr;
  TC_TRY(scratch_base(&base));
  Arena ar;
  ar.carve = true;
  ar.base = base;
  ar.cap = kNoAllocScratchBytes;
  Arena *prev = current_arena();
  current_arena() = &ar;
  const int rc = exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges,
                                  nullptr, nullptr, (unsigned long long *)count, params,
                                  (hipStream_t)stream);
  current_arena() = prev;
  TC_CHECK(rc != TC_OK || ar.total <= need, TC_ERR_HIP,
           "scratch use %zu exceeded its bound %zu", ar.total, need);
  return rc;
}

extern "C" int tc_edge_supports(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                                const int32_t *edge_u, const int32_t *edge_v, int64_t n_edges,
                                int32_t *out, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(row_ptr && out && n_edges >= 0, TC_ERR_INVALID, "invalid arguments");
  TC_CHECK(n_edges == 0 || (col_idx && edge_u && edge_v), TC_ERR_INVALID,
           "col_idx, edge_u and edge_v are required");
  TC_TRY(ensure_init());
  if (n_edges == 0) return TC_OK;
  hipStream_t s = (hipStream_t)stream;
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    hipLaunchKernelGGL(k_edge_supports<I>, dim3(grid_for(n_edges, kBlock, 65536)), dim3(kBlock),
                       0, s, Rows<I>{(const I *)row_ptr}, col_idx, edge_u, edge_v, n_edges, out);
  });
  TC_HIP(hipGetLastError());
  return TC_OK;
}
