/*
 * graph_utils_amd.h -- C ABI of the MI355X (gfx950) multi-source search path.
 *
 * These entry points replace graph-utils' host-side operator surface for the
 * multi-search family; each declaration cites the reference interface it
 * stands in for.  The C++ signatures of the reference (device_graph,
 * multi_search_*_setup, count_shortest_paths_setup,
 * betweenness_centrality_setup, diameter_sampling_setup, configure_grid,
 * transfer_result) are provided on top of this ABI by
 * graph-utils_amd/host/graph_utils.hpp.
 *
 * Conventions
 *  - Every call returns int status: GU_OK (0) or a negative GU_ERR_*; nothing
 *    exits the process (the reference's checkCudaErrors calls exit(),
 *    util_device.cuh:10-22).
 *  - Pointers named *_dev are device (HBM) pointers; all others are host.
 *    Host buffers are caller-owned and only borrowed for the call.
 *  - A gu_graph_t owns its device copy of the graph and a reusable search
 *    workspace; use one handle from one host thread at a time.
 *  - Result rows are source-major: row i of a [nsrc x n] output belongs to
 *    sources[i] (pitch n elements), as the reference's pitched rows.
 */
#ifndef GRAPH_UTILS_AMD_H
#define GRAPH_UTILS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GU_OK 0
#define GU_ERR_INVALID (-1)     /* bad argument (null pointer, size, option) */
#define GU_ERR_HIP (-2)         /* HIP runtime error */
#define GU_ERR_NOMEM (-3)       /* device or host allocation failed */
#define GU_ERR_RANGE (-4)       /* vertex id / source out of range */
#define GU_ERR_PARSE (-5)       /* malformed graph text */
#define GU_ERR_WEIGHTED (-6)    /* weighted METIS header (parse.cpp:164-173) */
#define GU_ERR_IO (-7)          /* file cannot be opened */
#define GU_ERR_NODEVICE (-8)    /* no HIP device */
#define GU_ERR_CAPACITY (-9)    /* output buffer too small / over budget */
#define GU_ERR_TIMEOUT (-10)    /* a wait ran out of time */
#define GU_ERR_UNSUPPORTED (-11) /* the host cannot provide it (no power readings) */

/* Expansion strategies for top-down levels.  Each mirrors one reference
 * kernel; all produce identical distances and path counts. */
#define GU_VARIANT_LINEAR 0        /* linear_with_atomics.cu:42-159 */
#define GU_VARIANT_EDGE_PARALLEL 1 /* edge_parallel.cu:37-92 (merge-path LBS) */
#define GU_VARIANT_WARP 2          /* warp_based.cu:47-190 */
#define GU_VARIANT_SCAN 3          /* scan_based.cu:43-194 */
#define GU_VARIANT_CTA_WARP 4      /* CTA_warp.cu:42-289 */
#define GU_VARIANT_SHUFFLE 5       /* shuffle_based.cuh:81-286 */
#define GU_VARIANT_AUTO 6          /* autotuner picks per graph */

#define GU_DIR_AUTO 0      /* switch top-down / bottom-up per level */
#define GU_DIR_TOPDOWN 1   /* frontier expansion only (the reference's form) */
#define GU_DIR_BOTTOMUP 2  /* pull from the frontier only */

/* Search engines.  BATCHED runs up to 256 sources per launch sequence as
 * lanes of per-vertex bit masks, one pass over the graph per level for all of
 * them (best when sources share levels: low-diameter graphs).  PER_SOURCE
 * runs one workgroup per source with the whole level loop inside one launch
 * (best for high-diameter, bounded-degree graphs).  AUTO picks from the
 * degree profile (PER_SOURCE when max degree <= 64 and mean degree <= 8).
 * Both engines run every job (search, path counts, Brandes, diameter). */
#define GU_ENGINE_AUTO 0
#define GU_ENGINE_BATCHED 1
#define GU_ENGINE_PER_SOURCE 2

typedef struct gu_graph *gu_graph_t;

typedef struct {
  int32_t variant;      /* GU_VARIANT_* (default AUTO) */
  int32_t direction;    /* GU_DIR_* (default AUTO) */
  int32_t batch_width;  /* sources per launch: 64, 128 or 256 (0 = auto) */
  int32_t wg_size;      /* threads per workgroup: 256, 512, 1024 (0 = auto) */
  int32_t alpha;        /* top-down -> bottom-up when frontier edges exceed
                           unvisited edges / alpha (0 = default) */
  int32_t beta;         /* bottom-up -> top-down when the frontier has fewer
                           than n / beta vertices (0 = default) */
  int32_t sync_every;   /* levels launched between host checks (0 = adaptive) */
  int32_t engine;       /* GU_ENGINE_* (default AUTO) */
  int32_t level_graphs; /* batched engine: 1 = replay each chunk of levels as
                           a HIP graph captured on the handle's private
                           stream; 0 = launch the kernels one by one.  A
                           capture in progress makes HIP refuse legacy
                           (NULL) stream work from other threads, so graphs
                           are opt-in. */
  int32_t source_queue; /* per-source engine (searches): 0 = auto (one grid
                           taking the call's sources from a device queue
                           when they exceed one launch of 256), 1 = always
                           the queue, 2 = launches of 256 sources */
  void *stream;         /* hipStream_t; NULL = the default stream */
} gu_config;

typedef struct {
  double device_ms;        /* device time of the call (HIP events) */
  int64_t teps_units;      /* sum over sources of edges scanned from reached
                              vertices (GTEPS numerator) */
  int64_t reached;         /* sum over sources of reached vertices */
  int64_t dag_edges;       /* sum over sources of shortest-path DAG edges
                              (entries v->w with d[w] = d[v]+1); counted
                              when path counts are computed */
  int32_t max_level;       /* deepest BFS level over all sources */
  int32_t batches;         /* launches of the batched engine */
  int32_t topdown_levels;  /* levels run top-down (summed over batches) */
  int32_t bottomup_levels; /* levels run bottom-up */
  int32_t sigma_overflow;  /* 1 if any path count wrapped modulo 2^64 */
  int32_t variant_used;    /* GU_VARIANT_* actually run */
  int32_t batch_width_used;
  int32_t engine_used;     /* GU_ENGINE_* actually run */
} gu_stats;

typedef struct {
  char name[256];
  char arch[64];           /* e.g. gfx950 */
  int32_t device;
  int32_t compute_units;   /* the reference's multiProcessorCount */
  int32_t max_threads_per_block;
  int32_t wave_size;
  int64_t total_mem;
  int64_t free_mem;
  double mem_bandwidth_gbs; /* bus width x memory clock x 2 */
  int32_t l2_bytes;
  int32_t clock_khz;
} gu_device_props;

/* ---- util_device ---------------------------------------------------- */
/* choose_device (util_device.cu:23-76): requested < 0 picks the device with
 * the highest memory bandwidth; fills *props for the chosen device. */
int gu_choose_device(int32_t requested, gu_device_props *props);
int gu_device_count(int32_t *count);
/* start_clock / end_clock (util_device.cu:4-21): HIP-event timer in ms on
 * the given stream. */
int gu_timer_start(void *stream, void **timer);
int gu_timer_stop(void *timer, double *ms);
/* power_measurer (util_device.cuh:47-62, util_device.cu:79-130): samples
 * the device's socket power through AMD SMI every period_ms on a host
 * thread between start and stop (nvmlDeviceGetPowerUsage there), and reads
 * the device's energy accumulator at both ends.  gu_power_start returns
 * GU_ERR_UNSUPPORTED where AMD SMI cannot be loaded or reports no power for
 * the device (the reference returns -1 on boards NVML cannot read).
 * gu_power_stop synchronizes the device first, as end_power_sample does,
 * fills *result and frees the handle. */
typedef struct gu_power_s *gu_power_t;
typedef struct {
  double avg_w;        /* mean of the samples, W (-1 without samples) */
  double max_w;        /* largest sample, W */
  double energy_avg_w; /* energy used / interval, W (-1 without the counter) */
  double energy_j;     /* energy used over the interval, J (-1 without it) */
  double seconds;      /* wall time from start to stop */
  int32_t samples;
} gu_power_result;
int gu_power_start(int32_t device, int32_t period_ms, gu_power_t *out);
int gu_power_stop(gu_power_t p, gu_power_result *result);
/* Trace ranges (no counterpart in the reference): roctxRangePushA /
 * roctxRangePop, so `rocprofv3 --marker-trace` attributes kernels to the
 * caller's phases.  The engine marks its own phases the same way.  Return
 * GU_ERR_UNSUPPORTED when the roctx library is absent (then a no-op). */
int gu_trace_push(const char *message);
int gu_trace_pop(void);
/* Memory helpers so host code needs no other runtime. */
int gu_malloc(void **p_dev, size_t bytes);
int gu_free(void *p_dev);
int gu_memcpy_h2d(void *dst_dev, const void *src, size_t bytes);
int gu_memcpy_d2h(void *dst, const void *src_dev, size_t bytes);
int gu_memset_d(void *dst_dev, int value, size_t bytes);
int gu_memcpy_d2d(void *dst_dev, const void *src_dev, size_t bytes);
int gu_synchronize(void);
const char *gu_strerror(int status);

/* ---- device_graph (device_graph.h:7-22, device_graph.cpp:3-14) ------- */
/* Copies R[n+1], C[m] and (optional) F[m] to HBM.  F may be NULL: it is then
 * rebuilt on the device from R.  Validates R monotone and C in range.
 * directed = 0 declares the CSR symmetric (every entry v->w has its w->v,
 * as parse_metis and parse_market produce); the engines then pull path
 * counts and frontier membership over the same adjacency.  directed = 1
 * builds the in-edge CSR for the pulls.  The arguments are checked on the
 * host before any device call: GU_ERR_INVALID for a bad shape or a CSR
 * declared undirected whose entries are not symmetric (for example a Matrix
 * Market file with more entries than its header, whose surplus
 * parse_market drops), GU_ERR_RANGE for a column id outside [0, n). */
int gu_graph_create(const int32_t *R, const int32_t *C, const int32_t *F,
                    int32_t n, int64_t m, int32_t directed, gu_graph_t *out);
int gu_graph_destroy(gu_graph_t g);
/* Device pointers and sizes of a resident graph. */
int gu_graph_arrays(gu_graph_t g, const int32_t **R_dev, const int32_t **C_dev,
                    const int32_t **F_dev, int32_t *n, int64_t *m);
/* Bytes of HBM the handle holds (graph + workspace). */
int gu_graph_device_bytes(gu_graph_t g, int64_t *bytes);
/* A second handle on the same resident graph (device_graph.h:7-22 has one
 * owner; this adds concurrent users): it reads g's device arrays and owns
 * its own search workspace, so two handles can run searches on two streams
 * at once.  Destroy every shared handle before g (gu_graph_destroy on g
 * returns GU_ERR_INVALID while shared handles exist). */
int gu_graph_share(gu_graph_t g, gu_graph_t *out);

/* ---- multi_search ---------------------------------------------------- */
/* Multi-source BFS.  Replaces multi_search_{linear_atomics,edge_parallel,
 * warp_based,scan_based,CTA_warp_based,shuffle_based}_setup
 * (linear_with_atomics.cuh:14, edge_parallel.cuh:13, warp_based.cuh:15,
 * scan_based.cuh:15, CTA_warp.cuh:15, shuffle_based.cuh:16) and
 * count_shortest_paths_setup (count_shortest_paths.cuh:13).
 * dist_dev: [nsrc x n] int32 (INT_MAX = unreachable), may be NULL.
 * sigma_dev: [nsrc x n] uint64 path counts mod 2^64, may be NULL. */
int gu_multi_search(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                    int32_t *dist_dev, uint64_t *sigma_dev,
                    const gu_config *cfg, gu_stats *stats);

/* gu_multi_search with a device row pointer per source (host arrays of nsrc
 * pointers to n-entry device rows; a NULL array means not wanted; NULL
 * entries skip distance rows but path counts need every row), plus an
 * optional host array ecc[nsrc] receiving each source's eccentricity (its
 * largest finite distance, diameter_sampling's per-source maximum).  With
 * both row arrays NULL only eccentricities are computed.  Lets one launch
 * serve sources of several queries whose outputs live apart. */
int gu_multi_search_rows(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                         int32_t *const *dist_rows_dev,
                         uint64_t *const *sigma_rows_dev, int32_t *ecc,
                         const gu_config *cfg, gu_stats *stats);

/* Brandes betweenness over the given sources.  Replaces
 * betweenness_centrality_setup (betweenness_centrality.cuh:14).
 * delta_dev: [nsrc x n] float dependency rows (the reference's delta_h),
 * may be NULL.  bc_dev: [n] double, BC(v) = sum_s delta_s(v), accumulated
 * into (caller zeroes it first), may be NULL. */
int gu_betweenness(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                   float *delta_dev, double *bc_dev, const gu_config *cfg,
                   gu_stats *stats);

/* gu_betweenness with per-source outputs: delta_rows_dev[i] (or NULL) gets
 * source i's dependency row; source i's dependencies are added into the
 * [n] double vector bc_targets_dev[i] (or not accumulated when NULL).
 * Within one launch, consecutive sources with the same target are summed
 * in source order and added once, so a query's sources packed together
 * accumulate as in gu_betweenness.  Two calls running at once must not
 * share a target. */
int gu_betweenness_rows(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                        float *const *delta_rows_dev,
                        double *const *bc_targets_dev, const gu_config *cfg,
                        gu_stats *stats);

/* Largest finite distance from any of the sources.  Replaces
 * diameter_sampling_setup (diameter_sampling.cuh:13). */
int gu_diameter(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                int32_t *diameter, const gu_config *cfg, gu_stats *stats);

/* ---- autotuner (no counterpart in the reference) ---------------------- */
typedef struct {
  int32_t engine;       /* GU_ENGINE_* chosen */
  int32_t variant;      /* GU_VARIANT_* (AUTO when the per-source engine won) */
  int32_t batch_width;  /* 0 = by source count */
  int32_t wg_size;      /* 0 = 256 */
  int32_t alpha;        /* 0 = default top-down -> bottom-up threshold */
  int32_t trials;       /* configurations timed */
  int32_t sources;      /* sources per trial */
  int32_t beta;         /* 0 = default bottom-up -> top-down threshold */
  double ms_default;    /* device ms of the untuned configuration */
  double ms_tuned;      /* device ms of the chosen one */
} gu_tune_result;

/* Times launches of the given sources (up to 256) for the job family `kind`
 * (GU_QUERY_BFS: searches with path counts, also used by diameter;
 * GU_QUERY_BC: Brandes) over engines, top-down expanders (shuffle, warp,
 * CTA+warp, scan, edge-parallel), sources per launch, workgroup sizes and
 * both direction-switch thresholds, starting from the degree-profile
 * defaults, and keeps the fastest in the handle: later calls of at least
 * half that many sources whose gu_config fields are left at auto / 0 use it
 * (handles made afterwards by gu_graph_share inherit it).  Allocates trial
 * outputs for the duration. */
int gu_autotune(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                int32_t kind, gu_tune_result *out);

/* ---- load_balanced_search (load_balanced_search.cuh:10-13) ----------- */
/* Device-wide form of extract_edges_block: scan_dev[i] = exclusive prefix of
 * counts_dev; result_dev[k] = upper_bound(scan, k) - 1 for k < W, the last
 * segment whose offset is <= k.  *W_out receives the total W.
 * Fails with GU_ERR_CAPACITY when W > result_cap (then nothing is written
 * to result_dev).  result_dev may be NULL: then only the scan and W are
 * computed (GU_OK).  Counts must be >= 0 and W < 2^31. */
int gu_lbs(const int32_t *counts_dev, int64_t S, int32_t *scan_dev,
           int32_t *result_dev, int64_t result_cap, int64_t *W_out,
           const gu_config *cfg, gu_stats *stats);

/* gu_lbs without any allocation, for callers that cannot allocate device
 * memory (code running under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  Same arguments, outputs and errors as gu_lbs, except:
 *  - its scratch (tile sums, the total, tile starts) is a fixed array of
 *    about 2.5 MiB in device memory that each GPU the library runs on holds
 *    until the process ends; the call makes no hipMalloc / hipFree and no
 *    HIP events, so stats->device_ms is 0;
 *  - S > GU_LBS_NOALLOC_MAX_SEGMENTS is refused with GU_ERR_CAPACITY before
 *    any device work; with a result buffer, W > GU_LBS_NOALLOC_MAX_ITEMS is
 *    refused with GU_ERR_CAPACITY after the scan (scan_dev and *W_out are
 *    written, result_dev is not);
 *  - calls on one GPU share its scratch: calls that overlap on one GPU
 *    (several streams or host threads at once) race and must be
 *    serialised by the caller. */
#define GU_LBS_NOALLOC_MAX_SEGMENTS (1LL << 27)
#define GU_LBS_NOALLOC_MAX_ITEMS (1LL << 28)
int gu_lbs_noalloc(const int32_t *counts_dev, int64_t S, int32_t *scan_dev,
                   int32_t *result_dev, int64_t result_cap, int64_t *W_out,
                   const gu_config *cfg, gu_stats *stats);

/* ---- query server (no counterpart in the reference) -------------------- */
/* A continuous-batching server over resident graphs.  Clients submit BFS,
 * Brandes BC and diameter-sampling queries of any number of sources from
 * any thread; the server admits them in arrival order while their outputs
 * fit an HBM budget, packs the sources of admitted queries of one graph and
 * kind into shared multi-source launches, and keeps up to `slots` launches
 * in flight (one stream and one shared graph handle per slot).  Outputs stay
 * in HBM in a block the query holds until gu_server_release.  A query's
 * distances, path counts and diameter are identical whatever it was packed
 * with; its BC differs only in the grouping of FP64 sums. */
typedef struct gu_server *gu_server_t;

#define GU_QUERY_BFS 0       /* distance rows (+ path-count rows) */
#define GU_QUERY_BC 1        /* BC vector over the query's sources */
#define GU_QUERY_DIAMETER 2  /* max eccentricity over the sources */
#define GU_QF_SIGMA 1        /* BFS: also path counts */
#define GU_QF_DELTA 2        /* BC: also per-source dependency rows */
#define GU_QF_HOST 4         /* deliver the outputs to host memory: the query
                                is done once they are in the server's pinned
                                host arena (gu_server_config.host_budget > 0) */
/* Result flag: a shortest-path count of one of the query's sources wrapped
 * modulo 2^64 (some reached vertex has at least 2^64 shortest paths, as on
 * large grids).  Distances stay exact and path counts exact mod 2^64, as the
 * reference's unsigned long long counts; dependencies and BC computed from
 * wrapped counts are not meaningful (they may be inf or NaN).  Exact per
 * source on the per-source engine; on the batched engine a launch in which
 * some source wrapped flags every query it carried. */
#define GU_QR_SIGMA_WRAPPED 1

typedef struct {
  int64_t hbm_budget;  /* bytes for query outputs (0 = a quarter of free HBM) */
  int32_t slots;       /* launches in flight (0 = 4) */
  int32_t max_pack;    /* sources per launch, <= 256 (0 = 256) */
  int32_t variant;     /* GU_VARIANT_* for every launch (0 = AUTO) */
  int32_t engine;      /* GU_ENGINE_* for every launch (0 = AUTO) */
  int64_t host_budget; /* bytes of pinned host memory for GU_QF_HOST query
                          outputs (0 = none: such queries are refused) */
} gu_server_config;

typedef struct {
  int32_t graph;            /* id from gu_server_add_graph */
  int32_t kind;             /* GU_QUERY_* */
  int32_t nsrc;             /* >= 1 */
  int32_t flags;            /* GU_QF_* */
  const int32_t *sources;   /* host, copied by gu_server_submit */
} gu_query;

typedef struct {
  int32_t status;           /* GU_OK or the launch's error */
  int32_t kind;
  int32_t nsrc;
  int32_t diameter;         /* GU_QUERY_DIAMETER */
  int32_t launches;         /* launches that carried its sources */
  int32_t flags;            /* GU_QR_* */
  int32_t *dist_dev;        /* [nsrc x n] (BFS) */
  uint64_t *sigma_dev;      /* [nsrc x n] (BFS with GU_QF_SIGMA) */
  float *delta_dev;         /* [nsrc x n] (BC with GU_QF_DELTA) */
  double *bc_dev;           /* [n] (BC) */
  double submit_ms;         /* server clock (ms since creation) */
  double admit_ms;
  double first_launch_ms;
  double done_ms;           /* with GU_QF_HOST: outputs in host memory */
  int64_t hbm_bytes;        /* arena bytes held until release */
  /* GU_QF_HOST: the outputs in the server's pinned host arena, valid until
   * gu_server_release (the *_dev pointers are then NULL: the query's device
   * block is given back as soon as the copy is done) */
  int32_t *dist_host;
  uint64_t *sigma_host;
  float *delta_host;
  double *bc_host;
} gu_query_result;

typedef struct {
  int64_t submitted, completed, rejected;
  int64_t launches, sources;
  int64_t hbm_in_use, hbm_peak, hbm_budget;
  int32_t waiting;          /* submitted, not yet admitted */
  int32_t active;           /* admitted, not finished */
  int64_t host_in_use, host_peak, host_budget;  /* pinned host arena */
} gu_server_stats;

/* Creates the server on the current device: the output arena, one stream
 * per slot and the worker threads. */
int gu_server_create(const gu_server_config *cfg, gu_server_t *out);
/* Registers a resident graph (the caller keeps ownership and destroys it
 * after the server) and allocates each slot's search workspace on it, so
 * that queries allocate no device memory; fails (GU_ERR_NOMEM) when the
 * device cannot hold that workspace. */
int gu_server_add_graph(gu_server_t s, gu_graph_t g, int32_t *graph_id);
/* Queues a query; returns at once.  A query whose outputs exceed the whole
 * budget finishes immediately with status GU_ERR_CAPACITY. */
int gu_server_submit(gu_server_t s, const gu_query *q, int64_t *ticket);
/* Waits for one query (timeout_ms < 0: no limit; GU_ERR_TIMEOUT). */
int gu_server_wait(gu_server_t s, int64_t ticket, int32_t timeout_ms,
                   gu_query_result *res);
/* Waits for the next finished query not yet reported. */
int gu_server_next_done(gu_server_t s, int32_t timeout_ms, int64_t *ticket,
                        gu_query_result *res);
int gu_server_poll(gu_server_t s, int64_t ticket, int32_t *done);
/* Frees a finished query's outputs (the device pointers become invalid). */
int gu_server_release(gu_server_t s, int64_t ticket);
int gu_server_now_ms(gu_server_t s, double *ms);
int gu_server_get_stats(gu_server_t s, gu_server_stats *out);
/* Stops the workers after their current launches and frees the server. */
int gu_server_destroy(gu_server_t s);

/* ---- host graph I/O and generators (parse.h:15-40) -------------------- */
/* Host arrays returned by these calls are released with gu_host_free. */
/* parse_metis (parse.cpp:115-208), including its tokenising behaviour:
 * the graph is undirected and m = 2 x header edge count. */
int gu_parse_metis(const char *path, int32_t **R, int32_t **C, int32_t **F,
                   int32_t *n, int64_t *m);
/* parse_market (parse.cpp:210-332): pattern Matrix Market, symmetrised
 * (every "u v" line adds both directions), entries sorted by (row, col). */
int gu_parse_market(const char *path, int32_t **R, int32_t **C, int32_t **F,
                    int32_t *n, int64_t *m);
/* parse_snap (parse.cpp:335-434): "u v" lines (0-based), '#' comments, a
 * directed graph with n = number of distinct ids; `header` is ignored as in
 * the reference (a header line is read as an entry). */
int gu_parse_snap(const char *path, int32_t header, int32_t **R, int32_t **C,
                  int32_t **F, int32_t *n, int64_t *m);
/* reverse (graph-utils/reverse.cpp:3-44): the transposed CSR, entries
 * ordered by (target, source); F2 = rows of the result. */
int gu_reverse(const int32_t *R, const int32_t *C, int32_t n, int64_t m,
               int32_t **R2, int32_t **C2, int32_t **F2);
/* host_graph::write_edgelist_to_file (parse.cpp:436-456). */
int gu_write_edgelist(const char *path, int32_t header, const int32_t *F,
                      const int32_t *C, int32_t n, int64_t m);
/* Writes a METIS file with single spaces and '\n' (1-based ids). */
int gu_write_metis(const char *path, const int32_t *R, const int32_t *C,
                   int32_t n, int64_t m);
/* Uniform simple undirected G(n, m_undirected): distinct pairs u<v drawn
 * without replacement, seeded. */
int gu_gen_er(int32_t n, int64_t m_undirected, uint64_t seed, int32_t **R,
              int32_t **C, int32_t **F, int32_t *n_out, int64_t *m_out);
/* R-MAT(scale, edge_factor, a, b, c), symmetrised, self loops and
 * duplicates removed, no label permutation, seeded. */
int gu_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                double c, uint64_t seed, int32_t **R, int32_t **C,
                int32_t **F, int32_t *n_out, int64_t *m_out);
/* side x side 4-neighbour grid, each edge deleted with probability p_delete,
 * then (if lcc_only) the largest component relabelled in increasing old id. */
int gu_gen_grid(int32_t side, double p_delete, uint64_t seed, int32_t lcc_only,
                int32_t **R, int32_t **C, int32_t **F, int32_t *n_out,
                int64_t *m_out);
/* Zipf(s) counts capped at cap, drawn until the sum reaches W, the last one
 * clipped so the sum is exactly W. */
int gu_gen_zipf_counts(double s, int32_t cap, int64_t W, uint64_t seed,
                       int32_t **counts, int64_t *S);
/* Seeded sample of k vertices with out-degree > 0 (with replacement). */
int gu_sample_sources(const int32_t *R, int32_t n, int32_t k, uint64_t seed,
                      int32_t *sources);
void gu_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
