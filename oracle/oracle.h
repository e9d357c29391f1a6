/*
 * oracle.h -- CPU restatement of graph-utils' sequential checkers.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is never linked into, called by or shipped with the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function names the reference file:line whose algorithm it restates.
 * Graph layout follows host_graph (parse.h:26-31): R[n+1] CSR offsets,
 * C[m] column ids, 0-based vertices; unreachable distance = INT_MAX.
 *
 * Pinning (see tests/test_oracle.py): every function here is checked against
 * closed forms (star, path, tree, grid binomials mod 2^64, tree BC formula),
 * brute-force path enumeration on tiny graphs, scipy shortest_path and
 * networkx betweenness on the reference's graphs/ fixtures, and numpy
 * searchsorted for the LBS.  No function is "parity unpinned".
 */
#ifndef GU_ORACLE_H
#define GU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FIFO BFS distances from one source.
 * examples/multi_search/multi_search_test.cu:17-38 */
void or_bfs(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
            int32_t *dist);

/* BFS distances + shortest-path counts (uint64, wraps mod 2^64).
 * examples/counting_shortest_paths/shortest_paths.cu:11-39 */
void or_sigma(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
              int32_t *dist, uint64_t *sigma);

/* Brandes dependency delta in float, push form over a LIFO stack of the BFS
 * pop order, skipping v == source.  examples/betweenness_centrality/bc.cu:11-57
 * dist / sigma may be NULL (scratch is allocated internally then). */
void or_brandes_f32(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
                    int32_t *dist, uint64_t *sigma, float *delta);

/* Same recurrence in double precision (sigma converted from uint64 to double
 * for the ratio).  This is the FP64 Brandes used to state BC error. */
void or_brandes_f64(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
                    int32_t *dist, uint64_t *sigma, double *delta);

/* BC = sum over sources of delta_s (no 1/2 factor, no n/k rescale), FP64.
 * shuffle_based.cu:350-353 (bc += delta per source). */
void or_bc_f64(const int32_t *R, const int32_t *C, int32_t n,
               const int32_t *sources, int32_t nsrc, double *bc);

/* or_bc_f64 on nthreads threads: static partition of the source list,
 * per-thread FP64 sums added in thread order (deterministic). */
void or_bc_f64_mt(const int32_t *R, const int32_t *C, int32_t n,
                  const int32_t *sources, int32_t nsrc, int32_t nthreads,
                  double *bc);

/* Max finite distance from src (eccentricity inside src's component).
 * examples/diameter_sampling/diameter_sampling_test.cu:9-38 */
int32_t or_eccentricity(const int32_t *R, const int32_t *C, int32_t n,
                        int32_t src);

/* Exclusive scan of counts; returns total W.  lbs.cu:44 (thrust scan). */
int64_t or_exclusive_scan(const int32_t *counts, int64_t S, int32_t *scan);

/* Load-balanced search by merging [0,W) with the scan (upper bound - 1).
 * examples/load_balanced_search/lbs.cu:18-31 */
void or_lbs(int64_t W, const int32_t *scan, int64_t S, int32_t *result);

/* Rows of a [start,end) multi-search result: row r holds the last source s
 * with s = start + r (mod G), s < end.  Returns the row count
 * min(end-start, G).  multi_search_test.cu:43-63, common.cu:17-25 */
int32_t or_rows_for_range(int32_t start, int32_t end, int32_t G,
                          int32_t *row_source);

/* TEPS units of one traversal: sum of out-degrees of reached vertices. */
int64_t or_teps_units(const int32_t *R, int32_t n, const int32_t *dist);

/* CPU baseline: BFS + sigma for a list of sources on nthreads threads
 * (static partition of the list), each source exactly as or_sigma.
 * Returns total TEPS units; *seconds gets wall time (steady clock). */
int64_t or_sigma_batch_timed(const int32_t *R, const int32_t *C, int32_t n,
                             const int32_t *sources, int32_t nsrc,
                             int32_t nthreads, double *seconds);

/* CPU baseline for the standalone LBS: scan + merge, single thread. */
double or_lbs_timed(const int32_t *counts, int64_t S, int32_t *scan,
                    int32_t *result, int64_t W);


/* Rows of or_sigma (sigma != NULL) or or_bfs (sigma == NULL) for a list of
 * sources on nthreads threads, plus each source's eccentricity (largest
 * finite distance, diameter_sampling_test.cu:9-38) when ecc != NULL.
 * dist may be NULL when only eccentricities are wanted (then sigma must be
 * NULL too).  Rows are [nsrc][n]. */
void or_rows_mt(const int32_t *R, const int32_t *C, int32_t n,
                const int32_t *sources, int32_t nsrc, int32_t nthreads,
                int32_t *dist, uint64_t *sigma, int32_t *ecc);

/* Full-size check of multi-search rows without recomputing them.  RT/CT is
 * the in-edge CSR (the adjacency itself for undirected graphs).  Row k
 * (source sources[k]) passes when
 *   d[src] = 0 and sigma[src] = 1;
 *   every other vertex with a finite distance has d[v] = 1 + min of d over
 *   its in-neighbours, and sigma[v] = sum of sigma over the in-neighbours u
 *   with d[u] = d[v] - 1 (mod 2^64);
 *   every vertex with d = INT_MAX has only INT_MAX in-neighbours and
 *   sigma = 0.
 * These conditions have exactly one solution, the BFS distances and
 * shortest-path counts (induction over the distance), so a passing row is
 * the oracle's row.  sigma may be NULL (distances only).  Returns the number
 * of violating (row, vertex) pairs; *first_bad = row * n + vertex of one of
 * them, or -1. */
int64_t or_check_search_rows(const int32_t *RT, const int32_t *CT, int32_t n,
                             const int32_t *sources, int32_t nrows,
                             const int32_t *dist, const uint64_t *sigma,
                             int32_t nthreads, int64_t *first_bad);

/* 1 when some vertex reached from src has at least 2^64 shortest paths (a
 * path count of or_sigma wraps), else 0. */
int32_t or_sigma_wraps(const int32_t *R, const int32_t *C, int32_t n,
                       int32_t src);

#ifdef __cplusplus
}
#endif
#endif
