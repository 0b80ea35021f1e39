/*
 * CPU oracle and CPU timing baseline for exact triangle counting
 * (test infrastructure only: loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg, never by the product path).
 *
 * Restates the reference's Numba CPU exact path
 *   cpu_exp/triangle_hll_experiment.py:543-603
 * (count_intersection_binary_numba / count_intersection_numba /
 *  exact_forward_count_numba): a parallel loop over oriented edges, each edge
 * intersecting two sorted CSR rows by merge, or by binary search of the shorter
 * row in the longer one when one row is more than 16x the other.
 * Policy 1 restates src/exact.py:29-41 instead (merge iff long/short <= 8).
 * All policies return the same integer count; only the timing differs.
 *
 * Build: see oracle/Makefile (gcc -O3 -march=x86-64-v3 -fopenmp -shared).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int64_t bsearch_count(const int32_t *adj, int64_t s0, int64_t s1,
                             int64_t l0, int64_t l1)
{
    /* cpu_exp:544-559: lower_bound of each short-row element in the long row */
    int64_t c = 0;
    for (int64_t p = s0; p < s1; ++p) {
        int32_t x = adj[p];
        int64_t lo = l0, hi = l1;
        while (lo < hi) {
            int64_t mid = lo + ((hi - lo) >> 1);
            if (adj[mid] < x) lo = mid + 1; else hi = mid;
        }
        if (lo < l1 && adj[lo] == x) ++c;
    }
    return c;
}

static int64_t merge_count(const int32_t *adj, int64_t a0, int64_t a1,
                           int64_t b0, int64_t b1)
{
    /* cpu_exp:578-592 */
    int64_t c = 0;
    while (a0 < a1 && b0 < b1) {
        int32_t x = adj[a0], y = adj[b0];
        if (x == y) { ++c; ++a0; ++b0; }
        else if (x < y) ++a0;
        else ++b0;
    }
    return c;
}

static inline int64_t edge_support(const int64_t *off, const int32_t *adj,
                                   int64_t a, int64_t b, int policy)
{
    int64_t a0 = off[a], a1 = off[a + 1], b0 = off[b], b1 = off[b + 1];
    int64_t la = a1 - a0, lb = b1 - b0;
    if (la == 0 || lb == 0) return 0;
    if (policy == 0) {                     /* cpu_exp:573-576 (16x rule) */
        if (la * 16 < lb) return bsearch_count(adj, a0, a1, b0, b1);
        if (lb * 16 < la) return bsearch_count(adj, b0, b1, a0, a1);
        return merge_count(adj, a0, a1, b0, b1);
    }
    /* src/exact.py:34-41 (r = 8 rule) */
    if (la > lb) {
        int64_t t;
        t = a0; a0 = b0; b0 = t; t = a1; a1 = b1; b1 = t; t = la; la = lb; lb = t;
    }
    if ((double)lb / (double)la <= 8.0) return merge_count(adj, a0, a1, b0, b1);
    return bsearch_count(adj, a0, a1, b0, b1);
}

int tco_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sum over the given oriented edges of |N+(u) ∩ N+(v)|.  When eu == NULL the
 * edge list is the CSR's own COO (u-major).  nthreads <= 0 uses all threads. */
uint64_t tco_exact_count(int64_t n, const int64_t *row_ptr, const int32_t *col,
                         int64_t n_edges, const int32_t *eu, const int32_t *ev,
                         int policy, int nthreads)
{
    uint64_t total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
    if (eu) {
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : total) num_threads(nthreads)
        for (int64_t e = 0; e < n_edges; ++e)
            total += (uint64_t)edge_support(row_ptr, col, eu[e], ev[e], policy);
    } else {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) num_threads(nthreads)
        for (int64_t u = 0; u < n; ++u)
            for (int64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k)
                total += (uint64_t)edge_support(row_ptr, col, u, col[k], policy);
    }
    return total;
}

/* Per-edge supports (int32) for the given edge list. */
void tco_edge_supports(const int64_t *row_ptr, const int32_t *col, int64_t n_edges,
                       const int32_t *eu, const int32_t *ev, int32_t *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads)
    for (int64_t e = 0; e < n_edges; ++e)
        out[e] = (int32_t)edge_support(row_ptr, col, eu[e], ev[e], 0);
}

/* splitmix64 finaliser (src/hll.py:13-19). */
static inline uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* HLL registers of N+(u) for every u with slot[u] >= 0, written to
 * regs[slot[u] * 2^p ...] (src/hll.py:81-96; src/gpu_hll.py:47-81).
 * The caller zeroes regs. */
void tco_hll_registers(int64_t n, const int64_t *row_ptr, const int32_t *col, int p,
                       const int64_t *slot, uint8_t *regs, int nthreads)
{
    const int rem = 64 - p;
    const uint64_t mask = (rem == 64) ? ~0ULL : ((1ULL << rem) - 1ULL);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t u = 0; u < n; ++u) {
        if (slot[u] < 0) continue;
        uint8_t *r = regs + (size_t)slot[u] * ((size_t)1 << p);
        for (int64_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
            uint64_t h = mix64((uint64_t)(int64_t)col[k]);
            uint64_t j = h >> rem;
            uint64_t w = h & mask;
            int bl = 0;
            while (w >> bl) ++bl;                  /* bit_length(w) */
            int rho = (w == 0) ? rem + 1 : rem - bl + 1;
            if (rho > r[j]) r[j] = (uint8_t)rho;
        }
    }
}
