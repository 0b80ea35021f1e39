// Internal interface of the exact-count pipeline, shared by the hybrid and
// batched entry points.
#pragma once

#include "tc_common.hpp"

namespace tc {

// Hybrid routing taken in the exact pipeline's edge pass (src/gpu_hybrid.py:80-89
// rule): edge e goes to HLL -- use[e] = 1, no exact count -- iff it wants HLL
// (route[e] != 0 when route is given, else min(D(u), D(v)) >= gamma with D the
// degrees array, or the row length when degrees is null) and both endpoints
// are sketched (has == 1); otherwise use[e] = 0.  counters[0] += HLL edges,
// counters[1] += edges that wanted HLL without both sketches.
// With net.params != null the learned router decides instead: an edge is a
// candidate iff both endpoints are sketched and min(D(u), D(v)) >=
// net.min_cost_degree, and a candidate goes to HLL iff the network's BF16
// output on its six features (tc_router.hpp) is >= net.threshold;
// counters[3] += candidates.
struct RouterNet {
  const float *params;   // kParams float32 (device)
  const float *norm;     // 12 float32 (device)
  const float *card;     // float32 [n]: single-sketch estimate of each vertex
  const uint8_t *regs;   // sketch rows of 2^p bytes, 16-byte aligned
  const int32_t *slot;   // compact row of each vertex (null: row = vertex id)
  int p;
  int sample_bits;       // features read the first 2^sample_bits registers
  float threshold;
  int64_t min_cost_degree;
};

struct HybridRoute {
  const int32_t *degrees;
  const uint8_t *has;
  int64_t gamma;
  const uint8_t *route;
  uint8_t *use;
  unsigned long long *counters;
  RouterNet net;
};

// Planner choices for a batch of graphs (sorted plan; all 0 = the
// single-graph form).  The batched entry points derive them from a device
// check of the batch's layout (tc_batch.hip); each is valid only under the
// conditions stated.
struct BatchPlan {
  // radix-sort only the low sort_bits bits of the center: the input's
  // packed words never interleave two centers whose low bits agree (edges
  // grouped by graph, every graph at most 2^sort_bits vertices)
  int sort_bits = 0;
  // row starts relative to the center's row start (PackLayout::rel_bits)
  int rel_bits = 0;
  // > 0: graph ids are below 2^graph_bits and no center row is longer than a
  // wavefront table holds; the runs are then ordered by graph and the
  // intersection kernel takes them from one queue per XCD (consecutive
  // graphs stay in one L2)
  int graph_bits = 0;
};

// count[0] (or count[graph_of[center]] when graph_of != nullptr) +=
// sum over edges e with !(skip && skip[e]) of |N+(eu[e]) ∩ N+(ev[e])|.
// eu == nullptr means the CSR's own COO (ev is then ignored).  With hr (sorted
// plan only, skip must be null) the edges hr routes to HLL are left out and
// hr->use and hr->counters are written for every edge.
int exact_count_impl(const void *row_ptr, int row_ptr_64, const int32_t *col, int64_t n,
                     const int32_t *eu, const int32_t *ev, int64_t E, const uint8_t *skip,
                     const int32_t *graph_of, unsigned long long *count,
                     const tc_exact_params *params, hipStream_t s,
                     const HybridRoute *hr = nullptr, const BatchPlan *bp = nullptr);

// Longest center row a wavefront's hash table holds.
constexpr int kWaveTableMaxDeg = 1023;

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v);
__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v);

}  // namespace tc
