/*
 * oracle.c -- CPU restatement of graph-utils' sequential checkers.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of the
 * product.  Plain C99; built by oracle/Makefile into oracle/liboracle.so.
 * Compiled without fast-math and without FP contraction so float arithmetic
 * is exactly IEEE single precision, as in the reference's sequential code.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <limits.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* multi_search_test.cu:17-38: distances start at INT_MAX, source 0, FIFO
 * queue, first discovery sets d[w] = d[v] + 1. */
void or_bfs(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
            int32_t *dist) {
  int32_t *Q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; i++) dist[i] = INT_MAX;
  dist[src] = 0;
  int32_t head = 0, tail = 0;
  Q[tail++] = src;
  while (head < tail) {
    int32_t v = Q[head++];
    for (int32_t k = R[v]; k < R[v + 1]; k++) {
      int32_t w = C[k];
      if (dist[w] == INT_MAX) {
        dist[w] = dist[v] + 1;
        Q[tail++] = w;
      }
    }
  }
  free(Q);
}

/* Forward pass shared by sigma and Brandes: shortest_paths.cu:11-39 and
 * bc.cu:22-43.  The pop order is written to order[] (the stack S of bc.cu
 * is this order read backwards).  Returns the number of reached vertices. */
static int32_t forward(const int32_t *R, const int32_t *C, int32_t n,
                       int32_t src, int32_t *dist, uint64_t *sigma,
                       int32_t *order) {
  for (int32_t i = 0; i < n; i++) {
    dist[i] = INT_MAX;
    sigma[i] = 0;
  }
  dist[src] = 0;
  sigma[src] = 1;
  int32_t head = 0, tail = 0;
  order[tail++] = src;
  while (head < tail) {
    int32_t v = order[head++];
    for (int32_t k = R[v]; k < R[v + 1]; k++) {
      int32_t w = C[k];
      if (dist[w] == INT_MAX) {
        dist[w] = dist[v] + 1;
        order[tail++] = w;
      }
      if (dist[w] == dist[v] + 1) sigma[w] += sigma[v]; /* wraps mod 2^64 */
    }
  }
  return tail;
}

void or_sigma(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
              int32_t *dist, uint64_t *sigma) {
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  forward(R, C, n, src, dist, sigma, order);
  free(order);
}

/* bc.cu:45-55: pop w from the stack; for every neighbour v one level closer
 * to the source (and v != source): delta[v] += sigma[v]/sigma[w]*(1+delta[w]),
 * the ratio taken in float after converting both counts to float. */
void or_brandes_f32(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
                    int32_t *dist, uint64_t *sigma, float *delta) {
  int32_t *d = dist ? dist : (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  uint64_t *s = sigma ? sigma : (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t cnt = forward(R, C, n, src, d, s, order);
  for (int32_t i = 0; i < n; i++) delta[i] = 0.0f;
  for (int32_t t = cnt - 1; t >= 0; t--) {
    int32_t w = order[t];
    for (int32_t k = R[w]; k < R[w + 1]; k++) {
      int32_t v = C[k];
      if (d[v] == d[w] - 1 && v != src) {
        float ratio = (float)s[v] / (float)s[w];
        delta[v] += ratio * (1.0f + delta[w]);
      }
    }
  }
  free(order);
  if (!dist) free(d);
  if (!sigma) free(s);
}

void or_brandes_f64(const int32_t *R, const int32_t *C, int32_t n, int32_t src,
                    int32_t *dist, uint64_t *sigma, double *delta) {
  int32_t *d = dist ? dist : (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  uint64_t *s = sigma ? sigma : (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t cnt = forward(R, C, n, src, d, s, order);
  for (int32_t i = 0; i < n; i++) delta[i] = 0.0;
  for (int32_t t = cnt - 1; t >= 0; t--) {
    int32_t w = order[t];
    for (int32_t k = R[w]; k < R[w + 1]; k++) {
      int32_t v = C[k];
      if (d[v] == d[w] - 1 && v != src)
        delta[v] += ((double)s[v] / (double)s[w]) * (1.0 + delta[w]);
    }
  }
  free(order);
  if (!dist) free(d);
  if (!sigma) free(s);
}

void or_bc_f64(const int32_t *R, const int32_t *C, int32_t n,
               const int32_t *sources, int32_t nsrc, double *bc) {
  double *delta = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; i++) bc[i] = 0.0;
  for (int32_t k = 0; k < nsrc; k++) {
    or_brandes_f64(R, C, n, sources[k], NULL, NULL, delta);
    for (int32_t i = 0; i < n; i++) bc[i] += delta[i];
  }
  free(delta);
}

typedef struct {
  const int32_t *R, *C, *sources;
  int32_t n, lo, hi;
  double *bc;
} bc_arg;

static void *bc_worker(void *p) {
  bc_arg *a = (bc_arg *)p;
  double *delta = (double *)malloc(sizeof(double) * (size_t)(a->n > 0 ? a->n : 1));
  for (int32_t i = 0; i < a->n; i++) a->bc[i] = 0.0;
  for (int32_t k = a->lo; k < a->hi; k++) {
    or_brandes_f64(a->R, a->C, a->n, a->sources[k], NULL, NULL, delta);
    for (int32_t i = 0; i < a->n; i++) a->bc[i] += delta[i];
  }
  free(delta);
  return NULL;
}

void or_bc_f64_mt(const int32_t *R, const int32_t *C, int32_t n,
                  const int32_t *sources, int32_t nsrc, int32_t nthreads,
                  double *bc) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nsrc) nthreads = nsrc > 0 ? nsrc : 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  bc_arg *args = (bc_arg *)malloc(sizeof(bc_arg) * (size_t)nthreads);
  for (int32_t t = 0; t < nthreads; t++) {
    args[t].R = R;
    args[t].C = C;
    args[t].sources = sources;
    args[t].n = n;
    args[t].lo = (int32_t)((int64_t)nsrc * t / nthreads);
    args[t].hi = (int32_t)((int64_t)nsrc * (t + 1) / nthreads);
    args[t].bc = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    pthread_create(&th[t], NULL, bc_worker, &args[t]);
  }
  for (int32_t i = 0; i < n; i++) bc[i] = 0.0;
  for (int32_t t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    for (int32_t i = 0; i < n; i++) bc[i] += args[t].bc[i];
    free(args[t].bc);
  }
  free(th);
  free(args);
}

/* diameter_sampling_test.cu:9-38: BFS, then the largest finite distance. */
int32_t or_eccentricity(const int32_t *R, const int32_t *C, int32_t n,
                        int32_t src) {
  int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  or_bfs(R, C, n, src, d);
  int32_t mx = 0;
  for (int32_t i = 0; i < n; i++)
    if (d[i] != INT_MAX && d[i] > mx) mx = d[i];
  free(d);
  return mx;
}

int64_t or_exclusive_scan(const int32_t *counts, int64_t S, int32_t *scan) {
  int64_t acc = 0;
  for (int64_t i = 0; i < S; i++) {
    scan[i] = (int32_t)acc;
    acc += counts[i];
  }
  return acc;
}

/* lbs.cu:18-31: merge the item indices [0,W) against the scanned offsets;
 * an item is emitted (as the current segment bi-1) while ai < scan[bi]. */
void or_lbs(int64_t W, const int32_t *scan, int64_t S, int32_t *result) {
  int64_t ai = 0, bi = 0;
  while (ai < W || bi < S) {
    int p;
    if (bi >= S)
      p = 1;
    else if (ai >= W)
      p = 0;
    else
      p = ai < (int64_t)scan[bi];
    if (p)
      result[ai++] = (int32_t)(bi - 1);
    else
      ++bi;
  }
}

int32_t or_rows_for_range(int32_t start, int32_t end, int32_t G,
                          int32_t *row_source) {
  int32_t total = end - start;
  if (total <= 0) return 0;
  int32_t rows = total < G ? total : G;
  if (total <= G) {
    for (int32_t i = 0; i < rows; i++) row_source[i] = start + i;
  } else {
    int32_t earliest = end - rows;
    int32_t loc = total % rows;
    for (int32_t i = 0; i < rows; i++) row_source[(loc + i) % rows] = earliest + i;
  }
  return rows;
}

int64_t or_teps_units(const int32_t *R, int32_t n, const int32_t *dist) {
  int64_t u = 0;
  for (int32_t i = 0; i < n; i++)
    if (dist[i] != INT_MAX) u += R[i + 1] - R[i];
  return u;
}

typedef struct {
  const int32_t *R, *C;
  int32_t n;
  const int32_t *sources;
  int32_t lo, hi;
  int64_t units;
} batch_arg;

static void *batch_worker(void *p) {
  batch_arg *a = (batch_arg *)p;
  int32_t n = a->n;
  int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  uint64_t *s = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  int64_t u = 0;
  for (int32_t k = a->lo; k < a->hi; k++) {
    int32_t cnt = forward(a->R, a->C, n, a->sources[k], d, s, order);
    for (int32_t t = 0; t < cnt; t++) u += a->R[order[t] + 1] - a->R[order[t]];
  }
  a->units = u;
  free(d);
  free(s);
  free(order);
  return NULL;
}

int64_t or_sigma_batch_timed(const int32_t *R, const int32_t *C, int32_t n,
                             const int32_t *sources, int32_t nsrc,
                             int32_t nthreads, double *seconds) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nsrc) nthreads = nsrc > 0 ? nsrc : 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  batch_arg *args = (batch_arg *)malloc(sizeof(batch_arg) * (size_t)nthreads);
  double t0 = now_s();
  for (int32_t t = 0; t < nthreads; t++) {
    args[t].R = R;
    args[t].C = C;
    args[t].n = n;
    args[t].sources = sources;
    args[t].lo = (int32_t)((int64_t)nsrc * t / nthreads);
    args[t].hi = (int32_t)((int64_t)nsrc * (t + 1) / nthreads);
    args[t].units = 0;
    if (nthreads == 1)
      batch_worker(&args[t]);
    else
      pthread_create(&th[t], NULL, batch_worker, &args[t]);
  }
  int64_t units = 0;
  for (int32_t t = 0; t < nthreads; t++) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    units += args[t].units;
  }
  if (seconds) *seconds = now_s() - t0;
  free(th);
  free(args);
  return units;
}

double or_lbs_timed(const int32_t *counts, int64_t S, int32_t *scan,
                    int32_t *result, int64_t W) {
  double t0 = now_s();
  int64_t tot = or_exclusive_scan(counts, S, scan);
  or_lbs(tot < W ? tot : W, scan, S, result);
  return now_s() - t0;
}

/* ---- threaded helpers for full-size parity checks ----------------------- */

typedef struct {
  const int32_t *R, *C, *sources;
  int32_t n, lo, hi;
  int32_t *dist;     /* [nsrc][n] or NULL */
  uint64_t *sigma;   /* [nsrc][n] or NULL */
  int32_t *ecc;      /* [nsrc] or NULL */
} rows_arg;

static void *rows_worker(void *p) {
  rows_arg *a = (rows_arg *)p;
  size_t n = (size_t)a->n;
  int32_t *dtmp = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int32_t k = a->lo; k < a->hi; k++) {
    int32_t *d = a->dist ? a->dist + (size_t)k * n : dtmp;
    if (a->sigma)
      or_sigma(a->R, a->C, a->n, a->sources[k], d, a->sigma + (size_t)k * n);
    else
      or_bfs(a->R, a->C, a->n, a->sources[k], d);
    if (a->ecc) {
      int32_t mx = 0;
      for (size_t i = 0; i < n; i++)
        if (d[i] != INT_MAX && d[i] > mx) mx = d[i];
      a->ecc[k] = mx;
    }
  }
  free(dtmp);
  return NULL;
}

void or_rows_mt(const int32_t *R, const int32_t *C, int32_t n,
                const int32_t *sources, int32_t nsrc, int32_t nthreads,
                int32_t *dist, uint64_t *sigma, int32_t *ecc) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nsrc) nthreads = nsrc > 0 ? nsrc : 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  rows_arg *args = (rows_arg *)malloc(sizeof(rows_arg) * (size_t)nthreads);
  for (int32_t t = 0; t < nthreads; t++) {
    rows_arg x = {R, C, sources, n, (int32_t)((int64_t)nsrc * t / nthreads),
                  (int32_t)((int64_t)nsrc * (t + 1) / nthreads), dist, sigma, ecc};
    args[t] = x;
    pthread_create(&th[t], NULL, rows_worker, &args[t]);
  }
  for (int32_t t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(args);
}

/* One row of or_check_search_rows.  Returns the violations; *bad gets the
 * first offending vertex. */
static int64_t check_row(const int32_t *RT, const int32_t *CT, int32_t n,
                         int32_t src, const int32_t *d, const uint64_t *s,
                         int64_t *bad) {
  int64_t viol = 0;
  *bad = -1;
#define GU_BAD(v)               \
  do {                          \
    if (*bad < 0) *bad = (v);   \
    viol++;                     \
  } while (0)
  if (src < 0 || src >= n) {
    GU_BAD(0);
    return viol;
  }
  for (int32_t v = 0; v < n; v++) {
    const int32_t dv = d[v];
    if (v == src) {
      if (dv != 0 || (s && s[v] != 1)) GU_BAD(v);
      continue;
    }
    int32_t mn = INT_MAX;
    uint64_t acc = 0;
    for (int32_t k = RT[v]; k < RT[v + 1]; k++) {
      const int32_t du = d[CT[k]];
      if (du < mn) mn = du;
      if (s && dv != INT_MAX && du == dv - 1) acc += s[CT[k]]; /* mod 2^64 */
    }
    if (dv == INT_MAX) {
      /* unreached: no in-neighbour reached, no paths */
      if (mn != INT_MAX || (s && s[v] != 0)) GU_BAD(v);
    } else {
      /* reached: one more than the nearest in-neighbour */
      if (dv < 1 || mn == INT_MAX || mn != dv - 1 || (s && s[v] != acc)) GU_BAD(v);
    }
  }
#undef GU_BAD
  return viol;
}

typedef struct {
  const int32_t *RT, *CT, *sources, *dist;
  const uint64_t *sigma;
  int32_t n, lo, hi;
  int64_t viol, bad;
} check_arg;

static void *check_worker(void *p) {
  check_arg *a = (check_arg *)p;
  size_t n = (size_t)a->n;
  a->viol = 0;
  a->bad = -1;
  for (int32_t k = a->lo; k < a->hi; k++) {
    int64_t b;
    int64_t v = check_row(a->RT, a->CT, a->n, a->sources[k], a->dist + (size_t)k * n,
                          a->sigma ? a->sigma + (size_t)k * n : NULL, &b);
    if (v && a->bad < 0) a->bad = (int64_t)k * (int64_t)n + b;
    a->viol += v;
  }
  return NULL;
}

int64_t or_check_search_rows(const int32_t *RT, const int32_t *CT, int32_t n,
                             const int32_t *sources, int32_t nrows,
                             const int32_t *dist, const uint64_t *sigma,
                             int32_t nthreads, int64_t *first_bad) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nrows) nthreads = nrows > 0 ? nrows : 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  check_arg *args = (check_arg *)malloc(sizeof(check_arg) * (size_t)nthreads);
  for (int32_t t = 0; t < nthreads; t++) {
    check_arg x = {RT, CT, sources, dist, sigma, n,
                   (int32_t)((int64_t)nrows * t / nthreads),
                   (int32_t)((int64_t)nrows * (t + 1) / nthreads), 0, -1};
    args[t] = x;
    pthread_create(&th[t], NULL, check_worker, &args[t]);
  }
  int64_t viol = 0, bad = -1;
  for (int32_t t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    viol += args[t].viol;
    if (bad < 0 && args[t].bad >= 0) bad = args[t].bad;
  }
  if (first_bad) *first_bad = bad;
  free(th);
  free(args);
  return viol;
}

/* shortest_paths.cu:11-39 with the additions checked for a carry: returns 1
 * when some reached vertex has at least 2^64 shortest paths from src (the
 * first such vertex, in BFS order, has exact parent counts whose sum carries
 * out of 64 bits), else 0. */
int32_t or_sigma_wraps(const int32_t *R, const int32_t *C, int32_t n,
                       int32_t src) {
  size_t nn = (size_t)(n > 0 ? n : 1);
  int32_t *d = (int32_t *)malloc(sizeof(int32_t) * nn);
  uint64_t *s = (uint64_t *)malloc(sizeof(uint64_t) * nn);
  int32_t *Q = (int32_t *)malloc(sizeof(int32_t) * nn);
  for (int32_t i = 0; i < n; i++) {
    d[i] = INT_MAX;
    s[i] = 0;
  }
  d[src] = 0;
  s[src] = 1;
  int32_t head = 0, tail = 0, wrapped = 0;
  Q[tail++] = src;
  while (head < tail) {
    int32_t v = Q[head++];
    for (int32_t k = R[v]; k < R[v + 1]; k++) {
      int32_t w = C[k];
      if (d[w] == INT_MAX) {
        d[w] = d[v] + 1;
        Q[tail++] = w;
      }
      if (d[w] == d[v] + 1) {
        uint64_t t = s[w] + s[v];
        wrapped |= t < s[v];
        s[w] = t;
      }
    }
  }
  free(d);
  free(s);
  free(Q);
  return wrapped;
}
