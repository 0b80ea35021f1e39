// Edge-list text ingest on gfx950: SNAP-style "u v" lines -> raw int32 pairs.
//
// Line rules (those of the reference's streaming reader,
// experiment/exp_gpu.py:131-143, _iter_edges_raw):
//   * a line ends at '\n' or '\r' (universal newlines: "\r\n" is one break,
//     the empty line it would leave is skipped anyway);
//   * leading whitespace is skipped; an empty line, or one whose first byte is
//     a comment byte ('#' and '%' by default), is skipped;
//   * the line is split on whitespace; fewer than two fields -> skipped;
//   * the first two fields must be base-10 integers as Python's int() reads
//     them ([+-]digits, single '_' between digits); more fields are ignored;
//   * u == v (self-loop) -> skipped.
// A field that is not an integer is an error (the reference raises
// ValueError there).  Negative ids and ids above INT32_MAX are errors too:
// the reference stores ids in an int32 array (exp_gpu.py:300), where they
// would overflow or index from the end.
//
// Layout of the work: the text is cut into 4 KiB tiles, one workgroup per
// tile.  The workgroup stages its tile plus a 1 KiB halo in LDS with 16-byte
// loads; each thread owns 16 bytes and parses every line that STARTS in
// them (reading past the halo from global memory for very long lines).
// Pass 1 counts the pairs of each tile and records the first bad line; a
// scan over the tile counts gives each tile its output offset; pass 2
// parses again and writes the pairs in file order.
#include "tc_common.hpp"
#include "tc_exact.hpp"

namespace tc {

constexpr int kTileBytes = 4096;
constexpr int kSegBytes = kTileBytes / kBlock;  // bytes owned by one thread
constexpr int kHalo = 1024;                     // staged bytes past the tile
constexpr int kPre = 16;                        // staged bytes before the tile
constexpr int kStage = kPre + kTileBytes + kHalo;

enum : int {
  kLineSkip = 0,
  kLinePair = 1,
  kLineBadInt = 2,    // a field that int() rejects
  kLineNegative = 3,  // an id below zero
  kLineRange = 4,     // an id above INT32_MAX
};

// Up to 8 comment bytes packed little-endian; unused bytes are 0 (a C string
// cannot hold a NUL, so 0 never matches a real byte of the set).
struct CommentSet {
  uint64_t bytes;
  __device__ __forceinline__ bool has(int c) const {
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int b = (int)((bytes >> (8 * k)) & 0xffu);
      hit |= (b != 0) & (b == c);
    }
    return hit;
  }
};

__device__ __forceinline__ bool is_break(int c) { return c == '\n' || c == '\r'; }
// str.split() whitespace among single bytes, minus the two line breaks
__device__ __forceinline__ bool is_space(int c) {
  return c == ' ' || c == '\t' || c == 0x0b || c == 0x0c || (c >= 0x1c && c <= 0x1f);
}

// Byte sources for the line parser.  A source returns the byte at an index,
// '\n' past the end of the text (the end of the text ends a line as a break
// does), or -2 where it cannot see: the LDS stage past its staged bytes, after
// which the line is parsed again from global memory.  Staged bytes past the
// end of the text were filled with '\n'.
struct LdsBytes {
  const uint8_t *stage;  // the __shared__ stage
  __device__ __forceinline__ int operator()(int i) const {
    return i < kStage ? (int)stage[i] : -2;
  }
};
struct GlobalBytes {
  const uint8_t *text;
  int64_t n;
  __device__ __forceinline__ int operator()(int64_t p) const {
    return p < n ? (int)text[p] : '\n';
  }
};

enum : int { kLineOffStage = 7 };

// The line rules as one predicated state machine: each byte updates the phase
// and the current field with selects, so lanes that sit in different states
// of their own lines do not split into branches.  Phases: 0 leading space,
// 1 first field, 2 gap, 3 second field, 4 two fields read, 5 line skipped,
// 6 ran past the visible bytes.  Field values are exact below 10^18 (kFHuge
// beyond), enough to tell equal ids apart for the self-loop test.
enum : uint32_t { kFDig = 1, kFBad = 2, kFHuge = 4, kFUnder = 8, kFNeg = 16 };

template <class Src, class Index>
__device__ __forceinline__ int parse_line(const Src &at, Index i, const CommentSet &cs,
                                          int32_t &u, int32_t &v) {
  int phase = 0;
  uint32_t cf = 0, f0 = 0;  // flags of the current field and of the first field
  uint64_t cv = 0, v0 = 0;
  while (phase < 4) {
    const int c = at(i);
    ++i;
    const bool off = c < 0;
    const bool br = (c == '\n') | (c == '\r');
    const bool sp = (c == ' ') | (c == '\t') | (c == 0x0b) | (c == 0x0c) |
                    ((unsigned)(c - 0x1c) <= 3u);
    const bool tok = !off & !br & !sp;
    const bool start = tok & ((phase == 0) | (phase == 2));
    const bool comment = tok & (phase == 0) & cs.has(c);
    // the current field restarts with its first byte
    cf = start ? 0u : cf;
    cv = start ? 0ull : cv;
    // one byte of a field: [+-] at its start, digits, single '_' between digits
    const bool dig = (unsigned)(c - '0') <= 9u;
    const bool sign = start & ((c == '+') | (c == '-'));
    const bool under_ok = (c == '_') & ((cf & (kFDig | kFUnder)) == kFDig);
    const bool take = tok & dig & !(cf & kFHuge);
    const bool huge = take & (cv >= 100000000000000000ull);
    cv = (take & !huge) ? cv * 10u + (uint64_t)(c - '0') : cv;
    uint32_t nf = cf;
    nf |= (tok & dig) ? kFDig : 0u;
    nf |= (tok & !dig & !sign & !under_ok) ? kFBad : 0u;
    nf |= huge ? kFHuge : 0u;
    nf |= (start & (c == '-')) ? kFNeg : 0u;
    nf = (tok & dig) ? (nf & ~kFUnder) : nf;
    nf = under_ok ? (nf | kFUnder) : nf;
    cf = tok ? nf : cf;
    // the first field ends at the first space or break after it
    const bool end0 = (phase == 1) & !tok;
    f0 = end0 ? cf : f0;
    v0 = end0 ? cv : v0;
    int np = phase;
    np = (phase == 0) ? (br ? 5 : (comment ? 5 : (tok ? 1 : 0))) : np;
    np = (phase == 1) ? (br ? 5 : (sp ? 2 : 1)) : np;
    np = (phase == 2) ? (br ? 5 : (tok ? 3 : 2)) : np;
    np = (phase == 3) ? (tok ? 3 : 4) : np;
    phase = off ? 6 : np;
  }
  if (phase == 6) return kLineOffStage;
  if (phase == 5) return kLineSkip;
  // int() of both fields comes first (a field it rejects is an error whatever
  // the other holds); then u == v drops the line; then the range checks
  const uint32_t f1 = cf;
  const uint64_t v1 = cv;
  auto bad = [](uint32_t f) { return !(f & kFDig) || (f & (kFUnder | kFBad)); };
  if (bad(f0) || bad(f1)) return kLineBadInt;
  const bool z0 = v0 == 0 && !(f0 & kFHuge), z1 = v1 == 0 && !(f1 & kFHuge);
  const bool n0 = (f0 & kFNeg) && !z0, n1 = (f1 & kFNeg) && !z1;  // "-0" is 0
  if (!((f0 | f1) & kFHuge) && v0 == v1 && n0 == n1) return kLineSkip;  // self-loop
  if (n0 || n1) return kLineNegative;
  if ((f0 & kFHuge) || (f1 & kFHuge) || v0 > 0x7fffffffull || v1 > 0x7fffffffull)
    return kLineRange;
  u = (int32_t)v0;
  v = (int32_t)v1;
  return kLinePair;
}

// Lines longer than the staged halo (rare): parsed from global memory.
struct LineResult {
  int st;
  int32_t u, v;
};
__device__ __noinline__ LineResult parse_line_global(const uint8_t *text, int64_t n, int64_t p,
                                                     CommentSet cs) {
  LineResult r{0, 0, 0};
  r.st = parse_line(GlobalBytes{text, n}, p, cs, r.u, r.v);
  return r;
}

// Exclusive scan of one int per thread over the workgroup; *total gets the sum.
// Ends with a barrier, so wave_tot may be reused right after.
__device__ __forceinline__ int block_exclusive_scan(int x, int *wave_tot, int *total) {
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int inc = wave_inclusive_scan(x);
  if (lane == kWave - 1) wave_tot[w] = inc;
  __syncthreads();
  int before = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < kBlock / kWave; ++i) {
    const int t = wave_tot[i];
    before += (i < w) ? t : 0;
    sum += t;
  }
  *total = sum;
  __syncthreads();
  return before + inc - x;
}

// One workgroup per 4 KiB tile.  The tile's line starts are found bytewise
// (16 bytes per thread) and compacted into an LDS list; then thread j parses
// lines j, j + 256, ... so that every lane has a line of its own.  Pairs are
// written in line order: each round of 256 lines is ranked by a block scan.
template <bool kWrite>
__global__ void __launch_bounds__(kBlock)
k_parse_tiles(const uint8_t *__restrict__ text, int64_t n, int aligned, CommentSet cs,
              int32_t *__restrict__ tile_count, const int64_t *__restrict__ tile_off,
              int2 *__restrict__ pairs, unsigned long long *__restrict__ bad) {
  __shared__ __align__(16) uint8_t stage[kStage];
  __shared__ uint16_t starts[kTileBytes / 2];  // at most one line start per two bytes
  __shared__ int wave_tot[kBlock / kWave];
  const int64_t t0 = (int64_t)blockIdx.x * kTileBytes;
  const int64_t s0 = t0 - kPre;  // text offset of stage[0]
  // stage [t0 - kPre, t0 + kTileBytes + kHalo); outside the text reads as '\n'
  if (aligned) {
    uint4 *st4 = reinterpret_cast<uint4 *>(stage);
    for (int i = threadIdx.x; i < kStage / 16; i += kBlock) {
      const int64_t p = s0 + 16 * (int64_t)i;
      if (p >= 0 && p + 16 <= n) {
        st4[i] = *reinterpret_cast<const uint4 *>(text + p);
      } else {
        for (int b = 0; b < 16; ++b) {
          const int64_t q = p + b;
          stage[16 * i + b] = (q >= 0 && q < n) ? text[q] : (uint8_t)'\n';
        }
      }
    }
  } else {
    for (int i = threadIdx.x; i < kStage; i += kBlock) {
      const int64_t q = s0 + i;
      stage[i] = (q >= 0 && q < n) ? text[q] : (uint8_t)'\n';
    }
  }
  __syncthreads();

  // line starts in this thread's 16 bytes: a non-break byte after a break
  const int b0 = kPre + (int)threadIdx.x * kSegBytes;
  uint32_t mask = 0;
  int prev = stage[b0 - 1];
#pragma unroll
  for (int q = 0; q < kSegBytes; ++q) {
    const int c = stage[b0 + q];
    if (!is_break(c) && is_break(prev) && t0 + (int64_t)(b0 - kPre + q) < n) mask |= 1u << q;
    prev = c;
  }
  int n_lines;
  int at = block_exclusive_scan(__popc(mask), wave_tot, &n_lines);
  while (mask) {
    const int q = __ffs(mask) - 1;
    mask &= mask - 1;
    starts[at++] = (uint16_t)(b0 - kPre + q);
  }
  __syncthreads();

  int64_t out = kWrite ? tile_off[blockIdx.x] : 0;
  int cnt = 0;
  for (int r0 = 0; r0 < n_lines; r0 += kBlock) {  // uniform trip count
    const int k = r0 + (int)threadIdx.x;
    int st = kLineSkip;
    int32_t u = 0, v = 0;
    int64_t p = 0;
    if (k < n_lines) {
      const int i = kPre + starts[k];
      p = t0 + (i - kPre);
      st = parse_line(LdsBytes{stage}, i, cs, u, v);
      if (st == kLineOffStage) {
        const LineResult r = parse_line_global(text, n, p, cs);
        st = r.st;
        u = r.u;
        v = r.v;
      }
    }
    const int is_pair = st == kLinePair;
    if constexpr (kWrite) {
      int round;
      const int rank = block_exclusive_scan(is_pair, wave_tot, &round);
      if (is_pair) pairs[out + rank] = make_int2(u, v);
      out += round;
    } else {
      cnt += is_pair;
      if (st != kLinePair && st != kLineSkip)
        atomicMin(bad, ((unsigned long long)p << 3) | (unsigned long long)st);
    }
  }
  if constexpr (!kWrite) {
    int total;
    (void)block_exclusive_scan(cnt, wave_tot, &total);
    if (threadIdx.x == 0) tile_count[blockIdx.x] = total;
  }
}

}  // namespace tc

extern "C" int tc_parse_edge_list(const uint8_t *text, int64_t n_bytes, const char *comment_chars,
                                  int32_t *pairs, int64_t max_pairs, int64_t *n_pairs,
                                  int64_t *bad_offset, int *bad_kind, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_pairs, TC_ERR_INVALID, "n_pairs is required");
  TC_CHECK(n_bytes >= 0 && max_pairs >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK(n_bytes == 0 || text, TC_ERR_INVALID, "text is required");
  TC_CHECK(((uintptr_t)pairs & 7) == 0, TC_ERR_INVALID, "pairs must be 8-byte aligned");
  const int64_t n_tiles = (n_bytes + kTileBytes - 1) / kTileBytes;
  TC_CHECK(n_tiles < INT32_MAX, TC_ERR_INVALID, "text of %lld bytes is too long for one call",
           (long long)n_bytes);
  TC_TRY(ensure_init());
  *n_pairs = 0;
  if (bad_offset) *bad_offset = -1;
  if (bad_kind) *bad_kind = 0;
  if (n_bytes == 0) return TC_OK;
  hipStream_t s = (hipStream_t)stream;

  CommentSet cs{0};
  const char *cc = comment_chars ? comment_chars : "#%";
  TC_CHECK(strlen(cc) <= 8, TC_ERR_INVALID, "at most 8 comment characters");
  for (int k = 0; cc[k]; ++k) cs.bytes |= (uint64_t)(unsigned char)cc[k] << (8 * k);
  const int aligned = ((uintptr_t)text & 15) == 0;

  Scratch cnt, off, bad;
  TC_TRY(cnt.alloc(sizeof(int32_t) * n_tiles, s));
  TC_TRY(off.alloc(sizeof(int64_t) * (n_tiles + 1), s));
  TC_TRY(bad.alloc(sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, bad.as<unsigned long long>(), 1,
                     ~0ull);
  hipLaunchKernelGGL(k_parse_tiles<false>, dim3((unsigned)n_tiles), dim3(kBlock), 0, s, text,
                     n_bytes, aligned, cs, cnt.as<int32_t>(), (const int64_t *)nullptr,
                     (int2 *)nullptr, bad.as<unsigned long long>());
  TC_HIP(hipGetLastError());
  TC_TRY(exclusive_scan<int32_t>(cnt.as<int32_t>(), off.as<int64_t>(), n_tiles, nullptr, s));
  int64_t total = 0;
  unsigned long long first_bad = 0;
  if (hipMemcpyAsync(&total, off.as<int64_t>() + n_tiles, sizeof(int64_t), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipMemcpyAsync(&first_bad, bad.as<unsigned long long>(), sizeof(first_bad),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "edge-list count pass failed: %s",
              hipGetErrorString(hipGetLastError()));
    return TC_ERR_HIP;
  }
  if (first_bad != ~0ull) {
    const int64_t at = (int64_t)(first_bad >> 3);
    const int kind = (int)(first_bad & 7);
    if (bad_offset) *bad_offset = at;
    if (bad_kind) *bad_kind = kind;
    set_error(TC_ERR_INVALID, "line at byte %lld: %s", (long long)at,
              kind == kLineBadInt     ? "expected integer node ids"
              : kind == kLineNegative ? "negative node id"
                                      : "node id does not fit int32");
    return TC_ERR_INVALID;
  }
  *n_pairs = total;
  if (!pairs) return TC_OK;  // sizing query
  if (total > max_pairs) {
    set_error(TC_ERR_INVALID, "%lld pairs do not fit the %lld-pair output", (long long)total,
              (long long)max_pairs);
    return TC_ERR_INVALID;
  }
  if (total > 0) {
    hipLaunchKernelGGL(k_parse_tiles<true>, dim3((unsigned)n_tiles), dim3(kBlock), 0, s, text,
                       n_bytes, aligned, cs, cnt.as<int32_t>(), (const int64_t *)off.as<int64_t>(),
                       reinterpret_cast<int2 *>(pairs), (unsigned long long *)nullptr);
    TC_HIP(hipGetLastError());
  }
  return TC_OK;
}

// ---------------------------------------------------------------------------
// The inverse: raw pairs -> SNAP-style text ("u<sep>v<newline>" per pair).
// No reference counterpart; used to stage large synthetic edge lists in HBM
// and for encode -> parse round trips.
// ---------------------------------------------------------------------------
namespace tc {

struct Delims {
  uint8_t sep[4], nl[4];
  int n_sep, n_nl;
};

__device__ __forceinline__ int dec_width(uint32_t x) {
  int w = 1;
  while (x >= 10u) {
    x /= 10u;
    ++w;
  }
  return w;
}

__global__ void k_line_bytes(const int2 *__restrict__ pairs, int64_t m, Delims d,
                             int32_t *__restrict__ len, unsigned int *__restrict__ negative) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int2 e = pairs[i];
    if ((e.x | e.y) < 0) atomicOr(negative, 1u);
    len[i] = dec_width((uint32_t)e.x) + dec_width((uint32_t)e.y) + d.n_sep + d.n_nl;
  }
}

__device__ __forceinline__ int64_t put_dec(uint8_t *__restrict__ out, int64_t o, uint32_t x) {
  const int w = dec_width(x);
  for (int k = w - 1; k >= 0; --k) {
    out[o + k] = (uint8_t)('0' + x % 10u);
    x /= 10u;
  }
  return o + w;
}

__global__ void k_write_lines(const int2 *__restrict__ pairs, int64_t m, Delims d,
                              const int64_t *__restrict__ off, uint8_t *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int2 e = pairs[i];
    int64_t o = put_dec(out, off[i], (uint32_t)e.x);
    for (int j = 0; j < d.n_sep; ++j) out[o++] = d.sep[j];
    o = put_dec(out, o, (uint32_t)e.y);
    for (int j = 0; j < d.n_nl; ++j) out[o++] = d.nl[j];
  }
}

}  // namespace tc

extern "C" int tc_format_edge_list(const int32_t *pairs, int64_t m, const char *sep,
                                   const char *newline, uint8_t *text, int64_t max_bytes,
                                   int64_t *n_bytes, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_bytes, TC_ERR_INVALID, "n_bytes is required");
  TC_CHECK(m >= 0 && max_bytes >= 0, TC_ERR_INVALID, "negative size");
  TC_CHECK(m == 0 || pairs, TC_ERR_INVALID, "pairs is required");
  TC_CHECK(((uintptr_t)pairs & 7) == 0, TC_ERR_INVALID, "pairs must be 8-byte aligned");
  Delims d{};
  const char *sp = sep ? sep : "\t", *nl = newline ? newline : "\n";
  d.n_sep = (int)strlen(sp);
  d.n_nl = (int)strlen(nl);
  TC_CHECK(d.n_sep >= 1 && d.n_sep <= 4 && d.n_nl >= 1 && d.n_nl <= 4, TC_ERR_INVALID,
           "separator and newline must be 1 to 4 bytes");
  memcpy(d.sep, sp, d.n_sep);
  memcpy(d.nl, nl, d.n_nl);
  TC_TRY(ensure_init());
  *n_bytes = 0;
  if (m == 0) return TC_OK;
  hipStream_t s = (hipStream_t)stream;
  Scratch len, off, neg;
  TC_TRY(len.alloc(sizeof(int32_t) * m, s));
  TC_TRY(off.alloc(sizeof(int64_t) * (m + 1), s));
  TC_TRY(neg.alloc(sizeof(unsigned int), s));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, neg.as<int32_t>(), 1, 0);
  const int2 *pp = reinterpret_cast<const int2 *>(pairs);
  hipLaunchKernelGGL(k_line_bytes, dim3(grid_for(m)), dim3(kBlock), 0, s, pp, m, d,
                     len.as<int32_t>(), neg.as<unsigned int>());
  TC_HIP(hipGetLastError());
  TC_TRY(exclusive_scan<int32_t>(len.as<int32_t>(), off.as<int64_t>(), m, nullptr, s));
  int64_t total = 0;
  unsigned int negative = 0;
  if (hipMemcpyAsync(&total, off.as<int64_t>() + m, sizeof(int64_t), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(&negative, neg.as<unsigned int>(), sizeof(negative), hipMemcpyDeviceToHost,
                     s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "edge-list sizing pass failed: %s",
              hipGetErrorString(hipGetLastError()));
    return TC_ERR_HIP;
  }
  TC_CHECK(!negative, TC_ERR_INVALID, "pairs hold a negative id");
  *n_bytes = total;
  if (!text) return TC_OK;  // sizing query
  TC_CHECK(total <= max_bytes, TC_ERR_INVALID, "%lld bytes of text do not fit %lld",
           (long long)total, (long long)max_bytes);
  hipLaunchKernelGGL(k_write_lines, dim3(grid_for(m)), dim3(kBlock), 0, s, pp, m, d,
                     (const int64_t *)off.as<int64_t>(), text);
  TC_HIP(hipGetLastError());
  return TC_OK;
}
