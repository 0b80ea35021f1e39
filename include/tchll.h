/*
 * tchll.h — C ABI of the MI355X (gfx950) triangle-counting hot path.
 *
 * Plain pointers and sizes only.  Every array argument is a DEVICE pointer
 * unless its comment says "host".  `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Calls are asynchronous on `stream` unless noted;
 * results written to device memory are valid once the stream is synchronised.
 *
 * Every function returns TC_OK (0) or a TC_ERR_* code; tc_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Each declaration names the reference interface it replaces
 * (ChitandaEruQ/triangle_counting_hll, file:line).
 */
#ifndef TCHLL_H
#define TCHLL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_OK 0
#define TC_ERR_INVALID 1     /* bad argument        -> Python ValueError   */
#define TC_ERR_HIP 2         /* HIP runtime failure -> Python RuntimeError */
#define TC_ERR_OOM 3         /* allocation failure  -> Python MemoryError  */
#define TC_ERR_UNSUPPORTED 4 /* unsupported device  -> Python RuntimeError */

#define TC_ABI_VERSION 1

/* ---- library ----------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
const char *tc_last_error(void);
/* TC_ABI_VERSION of the loaded library. */
int tc_abi_version(void);
/* Scratch memory: an entry point that needs scratch takes it from a
 * workspace kept per (device, stream), which grows on first use and is
 * reused by every later call on that stream (the stream orders the calls), so
 * repeated calls allocate nothing.  Calls that overlap on one stream from
 * several host threads are not supported.  tc_release_workspaces
 * synchronises the device and frees every workspace; call it only while no
 * other call is in flight.  No reference counterpart (the reference allocates
 * its accumulators per call, src/gpu_hybrid.py:343, 390-391). */
int tc_release_workspaces(void);

/* Device facts (host outputs).  Replaces the compute-capability probe of
 * src/gpu_hybrid.py:27-39 (_warn_if_float64_atomic_slow). */
int tc_device_info(int *n_cu, int *wave_size, int64_t *lds_per_block,
                   int64_t *total_mem, char *name, int name_len);

/* Per-kernel timing for measurement.  While enabled, the pipelines record a
 * hipEvent pair on the call's stream around each launch of the timed stages
 * (TC_STAGE_*).  tc_timing_read synchronises the recorded events, writes
 * the summed milliseconds and the launch count of every stage into ms[] and
 * launches[] (host, TC_N_STAGES entries each) and clears the log.
 * No reference counterpart (the reference times host wall-clock around
 * cuda.synchronize(), src/gpu_hybrid.py:346-352). */
#define TC_STAGE_CLASSIFY 0   /* light-edge merge + center histogram  */
#define TC_STAGE_HEAVY 1      /* LDS-hash intersection kernel          */
#define TC_STAGE_HLL_BUILD 2  /* sketch build                          */
#define TC_STAGE_HLL_EDGES 3  /* per-edge union estimates              */
#define TC_N_STAGES 4
int tc_timing_enable(int on);
int tc_timing_read(double *ms, int64_t *launches);

/* ---- (a) CSR build and orientation ------------------------------------- */

/* Opaque state between the two phases of the device CSR build. */
typedef struct tc_csr_build tc_csr_build;

/* Phase 1 of the device CSR build from raw undirected pairs (int32 [m][2],
 * ids >= 0).  Drops self-loops, dedups canonical (min,max) pairs, relabels
 * densely by sorted unique endpoint, computes undirected degrees and orients
 * u->v iff (deg u, u) < (deg v, v).  Host outputs: node count, oriented edge
 * count.  Synchronises `stream`.
 * Replaces the vectorised loader experiment/exp_gpu.py:264-389 (_load_graph)
 * and src/graph.py:29-61 (orient_graph). */
int tc_csr_build_begin(const int32_t *pairs, int64_t m, void *stream,
                       tc_csr_build **out, int64_t *n_nodes, int64_t *n_edges);

/* Edge-list text -> raw pairs (int32 [m][2], file order, self-loops dropped),
 * the input of tc_csr_build_begin.  `text` (device, n_bytes) holds whole
 * lines of a SNAP-style edge list; `comment_chars` (host, NUL-terminated,
 * at most 8 bytes, NULL = "#%") lists the bytes that start a comment line.  Lines, fields and
 * skips follow experiment/exp_gpu.py:131-143 (_iter_edges_raw): breaks at
 * '\n' / '\r', leading whitespace skipped, comment and one-field lines
 * skipped, the first two whitespace-separated fields read as Python int(),
 * further fields ignored.  A field int() rejects, a negative id or one above
 * INT32_MAX is an error: TC_ERR_INVALID with *bad_offset (host, optional) =
 * byte offset of the first such line and *bad_kind (host, optional) = 2 (not
 * an integer), 3 (negative) or 4 (too large).  *n_pairs (host) = number of
 * pairs; with pairs == NULL only that count is computed.  Output larger than
 * max_pairs is an error.  Synchronises `stream` after the counting pass; the
 * pairs are written asynchronously.  At most (n_bytes + 1) / 4 pairs fit in
 * any text.  Replaces the two Python passes of experiment/exp_gpu.py:291-313. */
int tc_parse_edge_list(const uint8_t *text, int64_t n_bytes, const char *comment_chars,
                       int32_t *pairs, int64_t max_pairs, int64_t *n_pairs,
                       int64_t *bad_offset, int *bad_kind, void *stream);

/* Raw pairs (int32 [m][2], ids >= 0) -> edge-list text, one
 * "u<sep>v<newline>" line per pair in order (`sep`, `newline`: host strings
 * of 1 to 4 bytes; NULL = "\t" and "\n").  *n_bytes (host) = text length;
 * with text == NULL only that length is computed.  Synchronises `stream`
 * after the sizing pass.  No reference counterpart: it stages synthetic edge
 * lists in HBM for tc_parse_edge_list and its round-trip tests. */
int tc_format_edge_list(const int32_t *pairs, int64_t m, const char *sep, const char *newline,
                        uint8_t *text, int64_t max_bytes, int64_t *n_bytes, void *stream);

/* Phase 2: writes the forward CSR into caller-owned arrays:
 *   row_ptr  [n+1]  int32 (row_ptr_64 = 0) or int64 (row_ptr_64 = 1)
 *   col_idx  [E]    int32, rows sorted ascending
 *   degrees  [n]    int32 forward degree
 *   node_ids [n]    int64 original id of each dense id (sorted)
 *   edge_u, edge_v [E] int32 COO of the CSR (u-major)
 *   undirected_deg [n] int32 (may be NULL)
 * Same output as src/gpu_graph.py:6-41 (to_csr_numpy) on the oriented dict. */
int tc_csr_build_finish(tc_csr_build *b, void *row_ptr, int row_ptr_64,
                        int32_t *col_idx, int32_t *degrees, int64_t *node_ids,
                        int32_t *edge_u, int32_t *edge_v, int32_t *undirected_deg,
                        void *stream);

/* Releases the phase-1 state (safe on NULL). */
void tc_csr_build_free(tc_csr_build *b);

/* Degree-bucket sort: vertices grouped by ceil(log2(deg+ + 1)) (33 buckets),
 * ascending id inside a bucket.  bucket_start [34] int64, perm [n] int32.
 * No reference counterpart (the reference sorts edges by cost with argsort,
 * experiment1_gpu/gpu_experiment_fixed.py:406-417). */
int tc_degree_buckets(const void *row_ptr, int row_ptr_64, int64_t n_nodes,
                      int64_t *bucket_start, int32_t *perm, void *stream);

/* ---- (b) exact counting ------------------------------------------------- */

/* Tuning knobs of the exact pipeline (0 = library default). */
typedef struct tc_exact_params {
  int light_max;     /* edges with deg+u + deg+v <= light_max use a lane merge */
  int chunk;         /* others per (center, chunk) work item                   */
  int waves_per_cu;  /* persistent waves per CU for the heavy kernel           */
  int unroll;        /* streamed ids in flight per lane: 1, 2, 4, 8 or 16     */
  int plan;          /* TC_PLAN_SORTED (default) or TC_PLAN_ATOMIC             */
  int block_min_bits; /* centers with bit_length(deg+) in [block_min_bits, 12]
                         are taken by whole workgroups over one joined LDS
                         table: 8..13 (13 = none), 0 = 10 (deg+ >= 512)       */
  int pack_ids;      /* sorted plan: 1 = packed words carry the other's id;
                         0 = they carry its row start and length when the
                         fields fit (no row-offset reads in the intersection
                         kernel), else its id                                 */
  int wide_offsets;  /* 1 = the intersection kernels address ids with 64-bit
                         offsets even when every row lies within the first
                         2^30 ids (the form graphs above 2^30 oriented edges
                         take; a test knob); 0 = chosen from row_ptr[n]       */
  int trim_rows;     /* sorted plan, row layout: 1 = the edge pass cuts each
                         streamed row to its ids <= the center row's largest
                         id (an upper-bound search per heavy edge; the ids
                         above it cannot be common); 0 = whole rows           */
  int tail_units;    /* intersection kernels, unroll 4..16: 1 = each streamed
                         row's last len % 4 ids are read in one 16-byte load
                         with its last whole ids, the ids past the row
                         masked (one memory line read once; masking work per
                         load); 0 = read one by one before the row's units   */
  int item_order;    /* sorted plan, one graph: 1 = the work items of the
                         wavefront phase are handed out in order of the col
                         position of their middle streamed row, in one range
                         per XCD (items running at once stream nearby rows,
                         so each XCD's L2 serves more of them; 4 items per
                         ticket); 0 = longest center row first from one
                         queue.  Ignored for batches (ordered by graph).   */
} tc_exact_params;

/* How heavy edges are grouped by center before the intersection kernel:
 * a radix sort of packed (center, other) words, or an atomic counting sort. */
#define TC_PLAN_SORTED 0
#define TC_PLAN_ATOMIC 1

/* count[0] += sum over the given oriented edges (edge_u[e] -> edge_v[e]) of
 * |N+(u) ∩ N+(v)|, with N+(x) = col_idx[row_ptr[x] : row_ptr[x+1]] sorted.
 * `count` is a device uint64.  edge_u/edge_v may both be NULL: the edge list
 * is then the CSR's own COO, and n_edges must equal row_ptr[n_nodes] (the
 * pass reads only the positions below both, so a mismatch miscounts but
 * never reads outside the rows).  params may be NULL.  Stream-ordered, except that
 * for at most 2^24 edges outside stream capture the call synchronises the
 * stream once, after the merge pass, and returns early when no edge needs the
 * hash-probe path (small inputs are bound by launch latency); the same holds
 * for the exact part of tc_hybrid_count.  n_nodes and n_edges must be below
 * 2^31 - 1 (the planner keeps edge positions in 32-bit fields); larger inputs
 * return TC_ERR_INVALID before any launch.
 * Replaces src/gpu_hybrid.py:96-130 (exact_count_kernel) and :332-354
 * (gpu_exact_triangle_count). */
int tc_exact_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                   int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                   int64_t n_edges, uint64_t *count, const tc_exact_params *params,
                   void *stream);

/* tc_exact_count without any allocation: no reference counterpart; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * result and stream behaviour as tc_exact_count (sorted plan only), but the
 * pipeline's scratch is carved from one 128 MiB array the library holds on
 * each device it runs on until the process ends; a graph whose scratch would
 * not fit (about 5M edges at 4 edges per vertex) is rejected with
 * TC_ERR_INVALID before any device work.  Calls that overlap on one device
 * (other streams, other host threads) share that array and race. */
int tc_exact_count_noalloc(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                           int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                           int64_t n_edges, uint64_t *count, const tc_exact_params *params,
                           void *stream);

/* Per-edge supports: out[e] = |N+(edge_u[e]) ∩ N+(edge_v[e])| (int32).
 * Replaces the per-edge records of src/exact.py:87-91. */
int tc_edge_supports(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                     const int32_t *edge_u, const int32_t *edge_v, int64_t n_edges,
                     int32_t *out, void *stream);

/* Stream compaction: idx receives, in increasing order, every i < n with
 * flags[i] != 0, and count (device int64) their number.  The device form of
 * the prototype's edge partition cp.where(light_mask)[0]
 * (experiment1_gpu/gpu_experiment_fixed.py:470); the library's own planners,
 * CSR build and sketch lists use the same kernels.  n < 2^31. */
int tc_compact_flags(const uint8_t *flags, int64_t n, int32_t *idx, int64_t *count,
                     void *stream);

/* ---- (c) HLL sketch build ----------------------------------------------- */

/* out[i] = splitmix64 finaliser of x[i] (uint64).  Exposes the device hash of
 * src/gpu_hll.py:12-17 (= src/hll.py:13-19) for testing. */
int tc_hash64(const uint64_t *x, uint64_t *out, int64_t n, void *stream);

/* Hash and estimator variants of the HLL entry points (hash_mode):
 *   TC_HLL_GPU      h = splitmix64 finaliser(id), index = top p bits, rank
 *                   from the low 64 - p bits; estimator of src/hll.py:37-54
 *                   (LinearCounting, then the 2^64 large-range correction).
 *                   The reference's GPU and src/ path (src/hll.py:13-19, 81-91).
 *   TC_HLL_CPU_EXP  h = splitmix64 finaliser(id + seed + 0x9E3779B97F4A7C15),
 *                   index = low p bits, rank from the high 64 - p bits;
 *                   estimator returning LinearCounting or the 2^32 large-range
 *                   correction of the raw value (cpu_exp/triangle_hll_experiment.py:
 *                   347-398, 615-635).  `seed` is used by this mode only.
 *   TC_HLL_PROTO32  h = 32-bit mixer of (uint32)id ^ 0x9747B28C, index = top p
 *                   bits, rank from the low 32 - p bits of (h << p) | 1;
 *                   estimator with LinearCounting and no large-range step;
 *                   in tc_hybrid_count the HLL support is est(u) + est(v) -
 *                   est(u | v), raised to 0 when negative (counted in
 *                   stats[2]) and not capped above (experiment1_gpu/
 *                   gpu_experiment_fixed.py:131-166, 209-254). */
#define TC_HLL_GPU 0
#define TC_HLL_CPU_EXP 1
#define TC_HLL_PROTO32 2

/* Builds HLL registers of N+(u) for every u with deg+(u) >= theta and sets
 * has_sketch[u] (uint8, every u).  Row of u: regs + slot[u] * 2^p when `slot`
 * (int32 [n], -1 = no row) is given, else regs + u * 2^p (dense layout of
 * src/gpu_hll.py:69).  accumulate = 1 takes the max with the bytes already in
 * the row (the in-place semantics of src/gpu_hll.py:80-81); accumulate = 0
 * overwrites the row.  4 <= p <= 16.
 * Replaces src/gpu_hll.py:47-81 (build_sketches_kernel). */
int tc_hll_build(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                 int64_t n_nodes, int64_t theta, int p, uint8_t *regs,
                 const int32_t *slot, uint8_t *has_sketch, int accumulate,
                 int hash_mode, uint64_t seed, void *stream);

/* ---- (d) union / intersection estimation (FP64) -------------------------- */

/* out[i] = FP64 HLL estimate of max(row(us[i]), row(vs[i])) register-wise
 * (us[i] == vs[i] gives the single-sketch estimate).  Rows addressed as in
 * tc_hll_build.  Replaces src/gpu_hll.py:88-121 (_union_estimate_device) with
 * the FP64 estimator of src/hll.py:37-54, 112-120. */
int tc_union_estimate(const uint8_t *regs, const int32_t *slot, int p,
                      const int32_t *us, const int32_t *vs, int64_t k,
                      double *out, int hash_mode, void *stream);

/* ---- (e) hybrid router and count ----------------------------------------- */

#define TC_PART_BOTH 0   /* fused: every edge, routed    (hybrid_count_kernel)      */
#define TC_PART_EXACT 1  /* exact-routed edges only      (hybrid_exact_part_kernel) */
#define TC_PART_HLL 2    /* HLL-routed edges only        (hybrid_hll_part_kernel)   */

/* Hybrid count over the given oriented edges.  deg(x) = degrees[x] when
 * `degrees` (int32 [n_nodes]) is given, else the CSR row length (the
 * reference passes the forward degrees).  Edge e wants HLL iff
 * route[e] != 0 when `route` (uint8 [n_edges]) is given, else iff
 * min(deg u, deg v) >= gamma; it takes the HLL path iff it wants HLL and
 * has_sketch[u] == has_sketch[v] == 1, else the exact path.
 * HLL support: s = deg u + deg v - union_estimate, clamped to [0, min].
 * row_ptr and col_idx may be NULL for part == TC_PART_HLL (then `degrees` is
 * required), and regs/p are ignored for part == TC_PART_EXACT.
 * Accumulates (device): counts[0] += FP64 sum of the supports of the edges of
 * `part` (deterministic order), stats[0] += n_approx, stats[1] += n_exact,
 * stats[2] += n_clamped (int64), and, when non-NULL, fallback[0] += edges
 * that wanted HLL but lacked a sketch.  per_edge (double [n_edges], may be
 * NULL) receives each processed edge's support.
 * Replaces src/gpu_hybrid.py:80-89, 137-325, 357-431 and src/hybrid.py:48-166. */
int tc_hybrid_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                    int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                    int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                    const int32_t *slot, const uint8_t *has_sketch, int64_t gamma, int p,
                    const uint8_t *route, int part, double *counts, int64_t *stats,
                    int64_t *fallback, double *per_edge, int hash_mode, void *stream);

/* tc_hybrid_count with the exact part's tuning knobs (tc_exact_params, NULL =
 * the library defaults, which tc_hybrid_count uses).  No reference
 * counterpart (the reference's hybrid has no knobs): for callers that tuned
 * the exact count on a graph and want its hybrid on the same settings.  The
 * knobs change the exact part's speed, never its supports. */
int tc_hybrid_count_tuned(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                          int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                          int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                          const int32_t *slot, const uint8_t *has_sketch, int64_t gamma, int p,
                          const uint8_t *route, int part, double *counts, int64_t *stats,
                          int64_t *fallback, double *per_edge, int hash_mode,
                          const tc_exact_params *params, void *stream);

/* ---- learned exact-vs-sketch router -------------------------------------- */
/* No reference counterpart: the reference routes by a fixed degree threshold
 * (src/hybrid.py:107-112, src/gpu_hybrid.py:80-89); these entry points let a
 * small MLP make that decision per edge. */

/* Features of candidate edges cand[0..n_cand) (edge ids into edge_u/edge_v),
 * float32 [n_cand][6]: log2(1 + deg u), log2(1 + deg v), their difference,
 * log2(1 + card[u]), log2(1 + card[v]) (card: float32 per-vertex sketch
 * estimates) and log2(1 + sampled union), the HLL estimate over the first
 * 2^sample_bits registers of both rows merged, times 2^(p - sample_bits).
 * Rows addressed as in tc_hll_build. */
int tc_router_features(const int32_t *degrees, const int32_t *edge_u, const int32_t *edge_v,
                       const int32_t *cand, int64_t n_cand, const uint8_t *regs,
                       const int32_t *slot, int p, int sample_bits, const float *card,
                       float *feats, void *stream);

/* Number of float32 parameters of the 6 -> 64 -> 1 network (513):
 * W1[64][6], b1[64], w2[64], b2.  `norm` holds mean[6] then inv_std[6]. */
int tc_router_n_params(void);

/* FP32 forward: prob[i] = sigmoid(w2 . relu(W1 x_i + b1) + b2). */
int tc_mlp_forward(const float *params, const float *norm, const float *feats, int64_t n,
                   float *prob, void *stream);

/* Gradient (float32 [513]) and value (*loss, device float) of the mean
 * weighted binary cross-entropy -(w t log p + (1 - t) log(1 - p)) over n
 * samples with targets t in [0, 1] and w = pos_weight.  Deterministic. */
int tc_mlp_grad(const float *params, const float *norm, const float *feats, const float *target,
                int64_t n, float pos_weight, float *grad, float *loss, void *stream);

/* One Adam step (bias-corrected; step counts from 1) on params in place. */
int tc_mlp_adam(float *params, const float *grad, float *m, float *v, int n_params, int step,
                float lr, float beta1, float beta2, float eps, void *stream);

/* BF16 inference (weights, inputs and hidden activations rounded to bf16,
 * FP32 accumulation).  route[cand[i]] (or route[i] when cand is NULL) =
 * prob_i >= threshold; prob (may be NULL) receives the probabilities. */
int tc_mlp_route_bf16(const float *params, const float *norm, const float *feats, int64_t n,
                      float threshold, const int32_t *cand, uint8_t *route, float *prob,
                      void *stream);

/* The network and its inputs for tc_router_hybrid_count (device pointers). */
typedef struct tc_router_net {
  const float *params;      /* 513 float32 (layout of tc_router_n_params)        */
  const float *norm;        /* 12 float32: mean[6], inv_std[6]                   */
  const float *card;        /* float32 [n_nodes]: single-sketch estimate          */
  float threshold;          /* HLL iff the BF16 output >= threshold              */
  int sample_bits;          /* features read the first 2^sample_bits registers    */
  int64_t min_cost_degree;  /* candidates need min(deg u, deg v) >= this         */
} tc_router_net;

/* Hybrid count routed by the network inside the exact pipeline's edge pass.
 * Edge e is a candidate iff has_sketch[u] == has_sketch[v] == 1 and
 * min(deg u, deg v) >= net->min_cost_degree (deg as in tc_hybrid_count);
 * for each candidate the pass computes the six features of
 * tc_router_features and the BF16 output of tc_mlp_route_bf16, and sends the
 * edge to HLL iff the output is >= net->threshold (the same decisions those
 * two entry points give).  Every other edge is counted exactly.  Accumulates
 * counts and stats as tc_hybrid_count with part TC_PART_BOTH; n_candidates
 * (device int64, may be NULL) += candidates; use (uint8 [n_edges], may be
 * NULL) receives 1 for HLL-routed edges, else 0; per_edge as in
 * tc_hybrid_count.  Replaces the per-edge predicate of src/gpu_hybrid.py:80-89
 * (and src/hybrid.py:107-112) inside the counting pass. */
int tc_router_hybrid_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                           int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                           int64_t n_edges, const int32_t *degrees, const uint8_t *regs,
                           const int32_t *slot, const uint8_t *has_sketch, int p,
                           const tc_router_net *net, double *counts, int64_t *stats,
                           int64_t *n_candidates, uint8_t *use, double *per_edge, void *stream);

/* ---- batched counting of many graphs ------------------------------------- */
/* No reference counterpart: the reference counts one graph per process
 * (experiment/exp_gpu.py:632-935).  Graphs are packed into one CSR over
 * disjoint id ranges (built by tc_csr_build_* from the concatenated pairs). */

/* graph_of[v] = the g with id_off[g] <= node_ids[v] < id_off[g+1]
 * (id_off: int64 [n_graphs + 1], ascending). */
int tc_vertex_graph_ids(const int64_t *node_ids, int64_t n_nodes, const int64_t *id_off,
                        int n_graphs, int32_t *graph_of, void *stream);

/* counts[g] += triangles of graph g (uint64 [n_graphs]); same pipeline as
 * tc_exact_count, each triangle credited to graph_of[center].  When graph_of
 * is nondecreasing, no edge joins two graphs and the edges are grouped by
 * graph in that order (the packed CSR's order), the radix sort of the heavy
 * edges covers only the bit width of the largest graph; a device check of
 * those conditions synchronises `stream` once per call. */
int tc_batch_exact_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                         int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                         int64_t n_edges, const int32_t *graph_of, int n_graphs,
                         uint64_t *counts, const tc_exact_params *params, void *stream);

/* Captures the batched pipeline (clear counts, then tc_batch_exact_count)
 * into one HIP graph over a workspace allocated here; the arrays passed are
 * bound into the graph.  Runs the pipeline once eagerly to size the
 * workspace (host output *workspace_bytes, may be NULL).  Synchronises. */
typedef struct tc_batch_graph tc_batch_graph;
int tc_batch_graph_create(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                          int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                          int64_t n_edges, const int32_t *graph_of, int n_graphs,
                          uint64_t *counts, const tc_exact_params *params, void *stream,
                          tc_batch_graph **out, int64_t *workspace_bytes);
/* One launch of the captured graph on `stream`: counts[] = per-graph counts. */
int tc_batch_graph_launch(tc_batch_graph *g, void *stream);
void tc_batch_graph_free(tc_batch_graph *g);

/* Raw pairs of many R-MAT / Erdos-Renyi graphs from the counter-based stream
 * draw(key, i) = splitmix64(key + (i + 1) * 0x9E3779B97F4A7C15).  Graph g:
 * kind[g] 0 = ER with param[g] = n, 1 = R-MAT with param[g] = scale and
 * quadrant thresholds t_a <= t_ab <= t_abc on 32-bit uniforms; its pairs fill
 * pairs[pair_off[g] .. pair_off[g+1]) (int32 [2] each) with ids offset by
 * id_off[g].  No reference counterpart (the reference reads SNAP files). */
int tc_generate_pairs(const uint8_t *kind, const int32_t *param, const int64_t *id_off,
                      const int64_t *pair_off, const uint64_t *key, int n_graphs, uint32_t t_a,
                      uint32_t t_ab, uint32_t t_abc, int32_t *pairs, int64_t n_pairs,
                      void *stream);

/* ---- DOULION edge sampling ---------------------------------------------- */

/* keep[i] = (numpy PCG64 draw first + i, as random()) < q for i in [0, n):
 * the stream of numpy.random.default_rng(seed).random(...), whose initial
 * 128-bit state and increment the caller reads from numpy
 * (bit_generator.state).  Draw i uses the generator state after first + i + 1
 * steps of s <- s * 0x2360ED051FC65DA44385DF649FCCF645 + inc, output XSL-RR,
 * double (x >> 11) * 2^-53.  Replaces `rng.random(n_fwd) < q` of
 * experiment/exp_gpu.py:414 (_doulion_forward) and src/doulion.py:38. */
int tc_pcg64_below(uint64_t state_hi, uint64_t state_lo, uint64_t inc_hi, uint64_t inc_lo,
                   uint64_t first, int64_t n, double q, uint8_t *keep, void *stream);

/* Edge-induced subgraph of a forward CSR: keeps edge e (CSR order) iff
 * keep[e] != 0, rows stay sorted.  out_row_ptr [n+1] int64; out_col holds up
 * to n_edges ids; *n_kept (host) = kept edges.  Synchronises `stream`.
 * Replaces the sample-to-forward-adjacency step of experiment/exp_gpu.py:415-433
 * (the sample keeps the original orientation). */
int tc_edge_subgraph(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                     int64_t n_nodes, int64_t n_edges, const uint8_t *keep,
                     int64_t *out_row_ptr, int32_t *out_col, int64_t *n_kept, void *stream);

/* ---- degree statistics ----------------------------------------------------- */

/* degree[u] += 1 and degree[v] += 1 for every edge (u, v) = (src[e * stride],
 * dst[e * stride]), e < n_edges (stride 1: two arrays; stride 2 with
 * dst = src + 1: interleaved pairs).  Ids must lie in [0, n_slots), else
 * TC_ERR_INVALID (checked; synchronises `stream`).
 * Replaces degree_count_kernel, experiment/gpu_graph_stats.py:89-97. */
int tc_degree_count(const int32_t *src, const int32_t *dst, int64_t stride, int64_t n_edges,
                    int32_t *degree, int64_t n_slots, void *stream);

/* Over the positive entries of degree[n_slots]: sums (device int64[3]) =
 * {count, sum, sum of squares}, minmax (device int32[2]) = {max, min}; both
 * overwritten ({0, 0, 0} and {0, INT32_MAX} when no entry is positive).
 * Replaces degree_stats_kernel, experiment/gpu_graph_stats.py:100-117. */
int tc_degree_summary(const int32_t *degree, int64_t n_slots, int64_t *sums, int32_t *minmax,
                      void *stream);

/* Histograms of the positive entries of degree[n_slots] (device outputs,
 * overwritten): hist[d] (n_hist = max degree + 1 entries) and
 * log_hist[min(floor(log2 d), n_log_buckets - 1)] (n_log_buckets <= 64).
 * A degree >= n_hist is TC_ERR_INVALID (checked; synchronises `stream`).
 * Replaces degree_hist_kernel, experiment/gpu_graph_stats.py:120-140. */
int tc_degree_histogram(const int32_t *degree, int64_t n_slots, int64_t *hist, int64_t n_hist,
                        int64_t *log_hist, int n_log_buckets, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TCHLL_H */
