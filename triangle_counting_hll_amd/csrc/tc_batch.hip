// Batched exact counting of many graphs packed into one CSR.
//
// The graphs occupy disjoint id ranges, so one device CSR build over the
// concatenated pairs keeps every graph's vertices contiguous, and one exact
// pipeline counts them all; triangles are credited to the graph of their
// center vertex.  For repeated counts the whole pipeline (fill, classify,
// radix plan, work items, intersection kernel) is captured once into a HIP
// graph over a preallocated workspace and replayed with one launch.
#include <algorithm>

#include "tc_common.hpp"
#include "tc_exact.hpp"

namespace tc {

__global__ void k_graph_of(const int64_t *__restrict__ node_ids, int64_t n,
                           const int64_t *__restrict__ id_off, int G,
                           int32_t *__restrict__ graph_of) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t id = node_ids[v];
    int lo = 0, hi = G;  // id_off[lo] <= id < id_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (id_off[mid] <= id) lo = mid; else hi = mid;
    }
    graph_of[v] = lo;
  }
}

// Layout facts of a batch that let the planner narrow its packed words and
// schedule by graph.  out[0] = max over graphs of their vertex count,
// out[1] = max over graphs of the ids their rows span (row_ptr[last + 1] -
// row_ptr[first]), out[2] != 0 when the layout does not qualify (graph ids
// decrease along the vertices, row_ptr decreases, an edge joins two graphs,
// or the edges are not grouped by graph in nondecreasing order of
// graph_of[edge_u]), out[3] = the largest forward degree.
template <class I>
__global__ void k_batch_layout(Rows<I> R, const int32_t *__restrict__ graph_of, int64_t n,
                               const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                               int64_t E, unsigned long long *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long verts = 0, span = 0, dmax = 0;
  bool bad = false;
  for (int64_t v = t0; v < n; v += stride) {
    const int32_t g = graph_of[v];
    const int32_t gn = v + 1 < n ? graph_of[v + 1] : g;
    const int64_t d = R.end(v) - R.begin(v);
    bad |= gn < g || d < 0;
    dmax = max(dmax, (unsigned long long)max<int64_t>(d, 0));
    if (v + 1 == n || gn != g) {
      // last vertex of graph g: its first vertex by binary search (exact
      // when graph_of is nondecreasing, else `bad` discards the result)
      int64_t lo = 0, hi = v;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (graph_of[mid] < g) lo = mid + 1; else hi = mid;
      }
      verts = max(verts, (unsigned long long)(v - lo + 1));
      const int64_t sp = R.end(v) - R.begin(lo);
      span = max(span, (unsigned long long)max<int64_t>(sp, 0));
    }
  }
  for (int64_t e = t0; e < E; e += stride) {
    const int32_t g = graph_of[eu[e]];
    bad |= graph_of[ev[e]] != g || (e + 1 < E && graph_of[eu[e + 1]] < g);
  }
  if (verts) atomicMax(&out[0], verts);
  if (span) atomicMax(&out[1], span);
  if (bad) atomicOr(&out[2], 1ull);
  if (dmax) atomicMax(&out[3], dmax);
}

static int bit_width(unsigned long long x) {
  int b = 0;
  while (b < 63 && (1ull << b) <= x) ++b;
  return b;
}

// The BatchPlan of a batch (all 0 unless k_batch_layout's conditions hold):
//   sort_bits: the bit width of the largest graph's vertex count minus one
//     (every center's packed words are then consecutive among those whose
//     centers share those low bits);
//   rel_bits: one more than the bit width of the largest span of ids a
//     graph's rows cover (row starts relative to the center's), 0 when that
//     is not narrower than row_ptr[n];
//   graph_bits: the bit width of the largest graph id, when no forward
//     degree exceeds a wavefront table (kWaveTableMaxDeg).
// Synchronises s (nothing is checked, and the plan stays all 0, while s is
// being captured into a graph).
static int batch_plan(const void *row_ptr, int row_ptr_64, const int32_t *graph_of, int64_t n,
                      const int32_t *eu, const int32_t *ev, int64_t E, hipStream_t s,
                      BatchPlan *bp) {
  *bp = BatchPlan{};
  if (n <= 0 || E <= 0) return TC_OK;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  TC_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return TC_OK;
  Scratch out;
  TC_TRY(out.alloc(6 * sizeof(unsigned long long), s));
  unsigned long long *o = out.as<unsigned long long>();
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kWave), 0, s, o, 6, 0ull);
  TC_DISPATCH_ROWS(row_ptr_64, I, {
    hipLaunchKernelGGL(k_batch_layout<I>, dim3(grid_for(std::max(n, E), kBlock, 4096)),
                       dim3(kBlock), 0, s, Rows<I>{(const I *)row_ptr}, graph_of, n, eu, ev, E,
                       o);
    // out[4] = row_ptr[n], out[5] = graph_of[n - 1] (low bytes)
    TC_HIP(hipMemcpyAsync(o + 4, (const I *)row_ptr + n, sizeof(I), hipMemcpyDeviceToDevice, s));
  });
  TC_HIP(hipMemcpyAsync(o + 5, graph_of + (n - 1), sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  TC_HIP(hipGetLastError());
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
  TC_HIP(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, s));
  TC_HIP(hipStreamSynchronize(s));
  if (h[2] != 0 || h[0] == 0) return TC_OK;
  bp->sort_bits = std::max(1, bit_width(h[0] - 1));
  const int rb = bit_width(h[1]) + 1;
  const unsigned long long total = row_ptr_64 ? h[4] : (unsigned long long)(uint32_t)h[4];
  if (rb < std::max(1, bit_width(total))) bp->rel_bits = rb;
  const int32_t last_graph = (int32_t)(uint32_t)h[5];
  if (last_graph >= 0 && h[3] <= (unsigned long long)kWaveTableMaxDeg)
    bp->graph_bits = std::max(1, bit_width((unsigned long long)last_graph));
  return TC_OK;
}

}  // namespace tc

struct tc_batch_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *workspace = nullptr;
  size_t workspace_bytes = 0;
  ~tc_batch_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (workspace) (void)hipFree(workspace);
  }
};

extern "C" int tc_vertex_graph_ids(const int64_t *node_ids, int64_t n_nodes, const int64_t *id_off,
                                   int n_graphs, int32_t *graph_of, void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_CHECK(n_nodes >= 0 && n_graphs > 0 && id_off && (n_nodes == 0 || (node_ids && graph_of)),
           TC_ERR_INVALID, "invalid arguments to tc_vertex_graph_ids");
  TC_TRY(ensure_init());
  if (n_nodes == 0) return TC_OK;
  hipLaunchKernelGGL(k_graph_of, dim3(grid_for(n_nodes, kBlock, 16384)), dim3(kBlock), 0,
                     (hipStream_t)stream, node_ids, n_nodes, id_off, n_graphs, graph_of);
  TC_HIP(hipGetLastError());
  return TC_OK;
}

static int check_batch_args(const void *row_ptr, const int32_t *col_idx, int64_t n_nodes,
                            const int32_t *edge_u, const int32_t *edge_v, int64_t n_edges,
                            const int32_t *graph_of, int n_graphs, uint64_t *counts) {
  using namespace tc;
  TC_CHECK(row_ptr && graph_of && counts && n_graphs > 0 && n_nodes >= 0 && n_edges >= 0,
           TC_ERR_INVALID, "invalid arguments to the batched count");
  TC_CHECK(n_edges == 0 || (col_idx && edge_u && edge_v), TC_ERR_INVALID,
           "col_idx, edge_u and edge_v are required");
  TC_CHECK(n_nodes < INT32_MAX && n_edges < INT32_MAX, TC_ERR_INVALID,
           "node ids and edge positions must fit int32");
  return TC_OK;
}

extern "C" int tc_batch_exact_count(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                                    int64_t n_nodes, const int32_t *edge_u, const int32_t *edge_v,
                                    int64_t n_edges, const int32_t *graph_of, int n_graphs,
                                    uint64_t *counts, const tc_exact_params *params,
                                    void *stream) {
  using namespace tc;
  Frame _frame((hipStream_t)stream);
  TC_TRY(check_batch_args(row_ptr, col_idx, n_nodes, edge_u, edge_v, n_edges, graph_of, n_graphs,
                          counts));
  TC_TRY(ensure_init());
  BatchPlan bp;
  TC_TRY(batch_plan(row_ptr, row_ptr_64, graph_of, n_nodes, edge_u, edge_v, n_edges,
                    (hipStream_t)stream, &bp));
  return exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges, nullptr,
                          graph_of, (unsigned long long *)counts, params, (hipStream_t)stream,
                          nullptr, &bp);
}

extern "C" int tc_batch_graph_create(const void *row_ptr, int row_ptr_64, const int32_t *col_idx,
                                     int64_t n_nodes, const int32_t *edge_u,
                                     const int32_t *edge_v, int64_t n_edges,
                                     const int32_t *graph_of, int n_graphs, uint64_t *counts,
                                     const tc_exact_params *params, void *stream,
                                     tc_batch_graph **out, int64_t *workspace_bytes) {
  using namespace tc;
  TC_CHECK(out, TC_ERR_INVALID, "null output handle");
  *out = nullptr;
  TC_TRY(check_batch_args(row_ptr, col_idx, n_nodes, edge_u, edge_v, n_edges, graph_of, n_graphs,
                          counts));
  TC_TRY(ensure_init());
  hipStream_t s = (hipStream_t)stream;
  auto *h = new tc_batch_graph();
  auto fail = [&](int rc) {
    current_arena() = nullptr;
    delete h;
    return rc;
  };

  // 1. eager run that measures the workspace (its counts are overwritten by
  //    the first replay, which starts by clearing them)
  BatchPlan bp;
  {
    const int rcb = batch_plan(row_ptr, row_ptr_64, graph_of, n_nodes, edge_u, edge_v, n_edges,
                               s, &bp);
    if (rcb) return fail(rcb);
  }
  Arena measure;
  current_arena() = &measure;
  int rc = exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges,
                            nullptr, graph_of, (unsigned long long *)counts, params, s, nullptr,
                            &bp);
  current_arena() = nullptr;
  if (rc) return fail(rc);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error(TC_ERR_HIP, "measuring run failed: %s", hipGetErrorString(hipGetLastError()));
    return fail(TC_ERR_HIP);
  }
  h->workspace_bytes = std::max<size_t>(measure.total, 256);
  if (hipMalloc(&h->workspace, h->workspace_bytes) != hipSuccess) {
    set_error(TC_ERR_OOM, "cannot allocate a %zu-byte graph workspace", h->workspace_bytes);
    return fail(TC_ERR_OOM);
  }

  // 2. capture the same allocation sequence carved from the workspace
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    set_error(TC_ERR_HIP, "cannot create the capture stream");
    return fail(TC_ERR_HIP);
  }
  Arena carve;
  carve.carve = true;
  carve.base = (char *)h->workspace;
  carve.cap = h->workspace_bytes;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipStreamDestroy(cs);
    set_error(TC_ERR_HIP, "hipStreamBeginCapture failed");
    return fail(TC_ERR_HIP);
  }
  current_arena() = &carve;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n_graphs, kBlock, 1024)), dim3(kBlock), 0, cs,
                     (unsigned long long *)counts, (int64_t)n_graphs, 0ull);
  rc = exact_count_impl(row_ptr, row_ptr_64, col_idx, n_nodes, edge_u, edge_v, n_edges, nullptr,
                        graph_of, (unsigned long long *)counts, params, cs, nullptr, &bp);
  current_arena() = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return fail(rc);
  }
  if (ec != hipSuccess || !g) {
    set_error(TC_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    return fail(TC_ERR_HIP);
  }
  h->graph = g;
  if (hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    set_error(TC_ERR_HIP, "hipGraphInstantiate failed");
    return fail(TC_ERR_HIP);
  }
  if (workspace_bytes) *workspace_bytes = (int64_t)h->workspace_bytes;
  *out = h;
  return TC_OK;
}

extern "C" int tc_batch_graph_launch(tc_batch_graph *h, void *stream) {
  using namespace tc;
  TC_CHECK(h && h->exec, TC_ERR_INVALID, "invalid batch graph handle");
  TC_HIP(hipGraphLaunch(h->exec, (hipStream_t)stream));
  return TC_OK;
}

extern "C" void tc_batch_graph_free(tc_batch_graph *h) { delete h; }
