// gu_persource.hpp -- one workgroup per source, the whole level loop inside
// one launch.
//
// For high-diameter, bounded-degree graphs (grids, meshes, road-like graphs)
// the sources of a batch sit on disjoint frontier rings at every level, so
// sharing a level across sources buys nothing while thousands of levels each
// pay global passes.  Here a workgroup owns one source, as the reference's
// kernels do (one block per source, shuffle_based.cuh:104), and separates
// levels with workgroup barriers instead of kernel launches.  Distances and
// path counts go straight into the caller's rows; the frontier
// queues live in HBM per workgroup.
//
// Visit (count_shortest_paths' visit, shuffle_based.cu:130-141): a compare-
// and-swap claims an unvisited vertex for level L+1; every edge v->w with
// d[w] = L+1 then adds sigma[v] to sigma[w] atomically (exact, mod 2^64).
#pragma once
#include "gu_kernels.hpp"

namespace gu {

struct PsArgs {
  const int *R;
  const int *C;
  int n;
  const int *srcs;   // [nsrc]
  int *const *dist;  // [nsrc] distance rows of n entries
  u64 *const *sigma; // [nsrc] path-count rows, or null
  int *queues;       // [nsrc][2][n]
  int *depth;        // [nsrc] deepest level reached
  u64 *part;         // [nsrc][3] (TEPS units, reached vertices, DAG edges)
  int *ovf;          // set when a path count wrapped
};

__device__ __forceinline__ u64 ld_l2(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_l2(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

#ifndef GU_PS_PREREAD
#define GU_PS_PREREAD 1
#endif
#ifndef GU_PS_SPEC
#define GU_PS_SPEC 0
#endif
#ifndef GU_PS_NORET
#define GU_PS_NORET 0
#endif

// Visits for NV frontier vertices of one thread, B adjacency entries of
// each per step: all column loads, then all claims, then all path-count
// adds, so a step costs three memory round trips for up to NV*B edges.  A
// claim is a compare-and-swap INT_MAX -> L+1 whose return is the vertex's
// distance either way, so no separate read of d precedes it.  Every edge
// v->w with d[w] = L+1 adds sigma[v] to sigma[w] (count_shortest_paths'
// visit, shuffle_based.cu:130-141); claimed vertices are appended to Q2 at
// positions taken from the level's LDS counter, one atomic per thread and
// step.
template <bool SIGMA, int NV, int B>
__device__ __forceinline__ void ps_visit_multi(const int *C, int *d, u64 *sg, int (&r)[NV],
                                               const int (&re)[NV], const u64 (&sv)[NV], int L, int &ovf,
                                               u64 &dag, unsigned *cnt, int *Q2) {
  while (true) {
    bool any = false;
#pragma unroll
    for (int u = 0; u < NV; u++) any |= r[u] < re[u];
    if (!any) break;
    int w[NV][B], o[NV][B];
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int j = 0; j < B; j++) w[u][j] = r[u] + j < re[u] ? C[r[u] + j] : -1;
#if GU_PS_PREREAD
    // a plain read first: only unvisited vertices take the atomic
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int j = 0; j < B; j++) o[u][j] = w[u][j] >= 0 ? d[w[u][j]] : -1;
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int j = 0; j < B; j++)
        if (o[u][j] == INT_MAX) o[u][j] = atomicCAS(&d[w[u][j]], INT_MAX, L + 1);
#else
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int j = 0; j < B; j++) o[u][j] = w[u][j] >= 0 ? atomicCAS(&d[w[u][j]], INT_MAX, L + 1) : -1;
#endif
    unsigned claims = 0;
#pragma unroll
    for (int u = 0; u < NV; u++)
#pragma unroll
      for (int j = 0; j < B; j++) claims += o[u][j] == INT_MAX;
    if (claims) {
      unsigned pos = atomicAdd(cnt, claims);
#pragma unroll
      for (int u = 0; u < NV; u++)
#pragma unroll
        for (int j = 0; j < B; j++)
          if (o[u][j] == INT_MAX) Q2[pos++] = w[u][j];
    }
    if (SIGMA) {
      u64 old[NV][B];
#if GU_PS_NORET
#pragma unroll
      for (int u = 0; u < NV; u++)
#pragma unroll
        for (int j = 0; j < B; j++) {
          old[u][j] = 0ull;
          if (o[u][j] == INT_MAX || o[u][j] == L + 1) (void)atomicAdd(&sg[w[u][j]], sv[u]);
        }
#else
#pragma unroll
      for (int u = 0; u < NV; u++)
#pragma unroll
        for (int j = 0; j < B; j++)
          old[u][j] = (o[u][j] == INT_MAX || o[u][j] == L + 1) ? atomicAdd(&sg[w[u][j]], sv[u]) : 0ull;
#endif
#pragma unroll
      for (int u = 0; u < NV; u++)
#pragma unroll
        for (int j = 0; j < B; j++)
          if (o[u][j] == INT_MAX || o[u][j] == L + 1) {
            ovf |= old[u][j] + sv[u] < sv[u];
            dag++;
          }
    }
#pragma unroll
    for (int u = 0; u < NV; u++) r[u] += B;
  }
}

// Vertices of degree above this are strip-mined by a whole wave.
constexpr int kPsSerialDeg = 32;
#ifndef GU_PS_NV
#define GU_PS_NV 1
#endif
#ifndef GU_PS_B
#define GU_PS_B 8
#endif
constexpr int kPsNV = GU_PS_NV, kPsB = GU_PS_B;

// Expansion inside the workgroup: each thread takes kPsNV frontier vertices
// at a time (ps_visit_multi; one vertex and 8-entry steps measured fastest
// with 256 sources in flight: more requests per thread only add memory
// contention); adjacency lists longer than kPsSerialDeg are
// strip-mined by the wave, the lanes electing the highest lane that still
// has one (race_and_resolve, warp_based.cu:109-145).  Levels are separated
// by one workgroup barrier: the level counters rotate over three LDS words,
// the one two levels back being cleared for the next level.
template <bool SIGMA>
__global__ __launch_bounds__(1024) void k_persource(PsArgs a) {
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf;
  u64 dag = 0;
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = lane_id();
  const int s = a.srcs[blockIdx.x];
  const size_t row = (size_t)blockIdx.x * a.n;
  int *d = a.dist[blockIdx.x];
  u64 *sg = SIGMA ? a.sigma[blockIdx.x] : nullptr;
  int *Q = a.queues + 2 * row;
  int *Q2 = Q + a.n;
  for (int i = tid; i < a.n; i += T) {
    d[i] = INT_MAX;
    if (SIGMA) sg[i] = 0;
  }
  if (tid < 3) s_cnt[tid] = 0;
  if (tid == 0) s_ovf = 0;
  __syncthreads();
  if (tid == 0) {
    d[s] = 0;
    if (SIGMA) sg[s] = 1;
    Q[0] = s;
  }
  __syncthreads();
  unsigned qn = 1;
  int L = 0;
  int ovf = 0;
  u64 teps = 0, reached = 1;
  while (true) {
    unsigned *cnt = &s_cnt[L % 3];
    if (tid == 0) s_cnt[(L + 1) % 3] = 0;
    for (unsigned base = 0; base < qn; base += kPsNV * T) {
      int r[kPsNV], re[kPsNV], rb[kPsNV], rbe[kPsNV];
      u64 sv[kPsNV];
#pragma unroll
      for (int u = 0; u < kPsNV; u++) {
        const unsigned k = base + u * T + tid;
        const int v = k < qn ? Q[k] : -1;
        r[u] = v >= 0 ? a.R[v] : 0;
        re[u] = v >= 0 ? a.R[v + 1] : 0;
        sv[u] = (SIGMA && v >= 0) ? ld_l2(&sg[v]) : 0ull;
      }
#pragma unroll
      for (int u = 0; u < kPsNV; u++) {
        teps += (u64)(re[u] - r[u]);
        const bool big = re[u] - r[u] > kPsSerialDeg;
        rb[u] = big ? r[u] : 0;
        rbe[u] = big ? re[u] : 0;
        if (big) re[u] = r[u];
      }
      ps_visit_multi<SIGMA, kPsNV, kPsB>(a.C, d, sg, r, re, sv, L, ovf, dag, cnt, Q2);
      // long lists: the wave strip-mines the elected lane's list
      for (int u = 0; u < kPsNV; u++) {
        int rr = rb[u];
        const int rre = rbe[u];
        while (true) {
          const bool pend = rr < rre;
          if (!ballot(pend)) break;
          const int win = race_and_resolve_wave(pend);
          const int rw = readlane(rr, win), rew = readlane(rre, win);
          const u64 svw = SIGMA ? readlane64(sv[u], win) : 0ull;
          if (lane == win) rr = rre;
          for (int e0 = rw; e0 < rew; e0 += kWave) {
            const int ed = e0 + lane;
            int w = -1, o = -1;
            if (ed < rew) {
              w = a.C[ed];
              o = atomicCAS(&d[w], INT_MAX, L + 1);
            }
            const bool claimed = o == INT_MAX;
            uint64_t m = ballot(claimed);
            if (m) {
              const int leader = __ffsll((unsigned long long)m) - 1;
              unsigned b = 0;
              if (lane == leader) b = atomicAdd(cnt, (unsigned)__popcll(m));
              b = (unsigned)readlane((int)b, leader);
              if (claimed) Q2[b + rank_in(m)] = w;
            }
            if (SIGMA && (claimed || o == L + 1)) {
              const u64 old = atomicAdd(&sg[w], svw);
              ovf |= old + svw < svw;
              dag++;
            }
          }
        }
      }
    }
    __syncthreads();
    const unsigned q2n = *cnt;
    if (q2n == 0) break;
    reached += q2n;
    int *t = Q;
    Q = Q2;
    Q2 = t;
    qn = q2n;
    L++;
  }
  // workgroup sums (lane-parallel per wave, then per workgroup)
  for (int o = 32; o >= 1; o >>= 1) {
    teps += __shfl_xor(teps, o, kWave);
    dag += __shfl_xor(dag, o, kWave);
  }
  if (ovf) atomicOr(&s_ovf, 1);
  if (lane == 0) {
    s_red[wave_id()][0] = teps;
    s_red[wave_id()][1] = dag;
  }
  __syncthreads();
  if (tid == 0) {
    u64 t = 0, g = 0;
    for (int w = 0; w < T / kWave; w++) {
      t += s_red[w][0];
      g += s_red[w][1];
    }
    a.part[3 * blockIdx.x] = t;
    a.part[3 * blockIdx.x + 1] = reached;
    a.part[3 * blockIdx.x + 2] = g;
    a.depth[blockIdx.x] = L;
    if (s_ovf) atomicOr(a.ovf, 1);
  }
}

// Brandes on one workgroup per source (the dependency accumulation of
// bc.cu:11-57 in pull form, as the reference's GPU backward,
// shuffle_based.cu:217-441): the forward pass above with the BFS order kept
// as a stack S[] of levels (E[l] = start of level l in S), then the levels
// from the deepest down, each vertex v summing over successors w (d[w] =
// d[v]+1, adjacency order) (sigma_v / sigma_w)(1 + delta_w) in FP32 with a
// Neumaier compensated sum; delta of the source stays 0 (bc.cu:50).
// Distances and path counts were written with L2 atomics, so the backward
// reads them through L2.
struct PsBcArgs {
  const int *R;
  const int *C;
  int n;
  const int *srcs;
  int *const *dist;     // [nsrc] scratch rows
  u64 *const *sigma;    // [nsrc] scratch rows
  float *const *delta;  // [nsrc] dependency rows
  int *stack;           // [nsrc][2n+2]: S, then level starts E
  int *depth;
  u64 *part;
  int *ovf;
};

__device__ __forceinline__ void ps_neumaier(float &sum, float &comp, float x) {
  float t = sum + x;
  if (fabsf(sum) >= fabsf(x))
    comp += (sum - t) + x;
  else
    comp += (x - t) + sum;
  sum = t;
}

__global__ __launch_bounds__(1024) void k_ps_bc(PsBcArgs a) {
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf;
  const int tid = threadIdx.x, T = blockDim.x;
  const int lane = lane_id();
  const int s = a.srcs[blockIdx.x];
  int *d = a.dist[blockIdx.x];
  u64 *sg = a.sigma[blockIdx.x];
  float *dl = a.delta[blockIdx.x];
  int *S = a.stack + (size_t)blockIdx.x * (2 * (size_t)a.n + 2);
  int *E = S + a.n;
  for (int i = tid; i < a.n; i += T) {
    d[i] = INT_MAX;
    sg[i] = 0;
    dl[i] = 0.0f;
  }
  if (tid < 3) s_cnt[tid] = 0;
  if (tid == 0) s_ovf = 0;
  __syncthreads();
  if (tid == 0) {
    d[s] = 0;
    sg[s] = 1;
    S[0] = s;
    E[0] = 0;
    E[1] = 1;
  }
  __syncthreads();
  unsigned qb = 0, qe = 1;
  int L = 0, ovf = 0;
  u64 teps = 0, reached = 1, dag = 0;
  while (true) {
    unsigned *cnt = &s_cnt[L % 3];
    if (tid == 0) s_cnt[(L + 1) % 3] = 0;
    for (unsigned base = qb; base < qe; base += kPsNV * T) {
      int r[kPsNV], re[kPsNV];
      u64 sv[kPsNV];
#pragma unroll
      for (int u = 0; u < kPsNV; u++) {
        const unsigned k = base + u * T + tid;
        const int v = k < qe ? S[k] : -1;
        r[u] = v >= 0 ? a.R[v] : 0;
        re[u] = v >= 0 ? a.R[v + 1] : 0;
        sv[u] = v >= 0 ? ld_l2(&sg[v]) : 0ull;
      }
#pragma unroll
      for (int u = 0; u < kPsNV; u++) teps += (u64)(re[u] - r[u]);
      ps_visit_multi<true, kPsNV, kPsB>(a.C, d, sg, r, re, sv, L, ovf, dag, cnt, S + qe);
    }
    __syncthreads();
    const unsigned q2n = *cnt;
    if (q2n == 0) break;
    reached += q2n;
    qb = qe;
    qe += q2n;
    L++;
    if (tid == 0) E[L + 1] = (int)qe;
  }
  __syncthreads();
  // backward, deepest level first (its vertices have no successors); each
  // thread takes kPsNV vertices, loads d, sigma and delta of a chunk's
  // neighbours together and sums the successors in adjacency order
  for (int l = L - 1; l >= 1; l--) {
    const int b0 = E[l], b1 = E[l + 1];
    for (int i0 = b0 + tid; i0 < b1; i0 += kPsNV * T) {
      int v[kPsNV], r[kPsNV], re[kPsNV];
      float sv[kPsNV], sum[kPsNV], comp[kPsNV];
#pragma unroll
      for (int u = 0; u < kPsNV; u++) {
        const int i = i0 + u * T;
        v[u] = i < b1 ? S[i] : -1;
        r[u] = v[u] >= 0 ? a.R[v[u]] : 0;
        re[u] = v[u] >= 0 ? a.R[v[u] + 1] : 0;
        sv[u] = v[u] >= 0 ? (float)ld_l2(&sg[v[u]]) : 0.0f;
        sum[u] = 0.0f;
        comp[u] = 0.0f;
      }
      while (true) {
        bool any = false;
#pragma unroll
        for (int u = 0; u < kPsNV; u++) any |= r[u] < re[u];
        if (!any) break;
        int w[kPsNV][kPsB];
        int dw[kPsNV][kPsB];
        float sw[kPsNV][kPsB], xw[kPsNV][kPsB];
#pragma unroll
        for (int u = 0; u < kPsNV; u++)
#pragma unroll
          for (int j = 0; j < kPsB; j++) w[u][j] = r[u] + j < re[u] ? a.C[r[u] + j] : -1;
#pragma unroll
        for (int u = 0; u < kPsNV; u++)
#pragma unroll
          for (int j = 0; j < kPsB; j++) {
            const int x = w[u][j] >= 0 ? w[u][j] : 0;
            dw[u][j] = w[u][j] >= 0 ? ld_l2(&d[x]) : -1;
#if GU_PS_SPEC
            sw[u][j] = (float)ld_l2(&sg[x]);
            xw[u][j] = dl[x];
#endif
          }
#if !GU_PS_SPEC
#pragma unroll
        for (int u = 0; u < kPsNV; u++)
#pragma unroll
          for (int j = 0; j < kPsB; j++) {
            const bool on = dw[u][j] == l + 1;
            sw[u][j] = on ? (float)ld_l2(&sg[w[u][j]]) : 1.0f;
            xw[u][j] = on ? dl[w[u][j]] : 0.0f;
          }
#endif
#pragma unroll
        for (int u = 0; u < kPsNV; u++)
#pragma unroll
          for (int j = 0; j < kPsB; j++)
            if (dw[u][j] == l + 1) ps_neumaier(sum[u], comp[u], (sv[u] / sw[u][j]) * (1.0f + xw[u][j]));
#pragma unroll
        for (int u = 0; u < kPsNV; u++) r[u] += kPsB;
      }
#pragma unroll
      for (int u = 0; u < kPsNV; u++)
        if (v[u] >= 0) dl[v[u]] = sum[u] + comp[u];
    }
    __syncthreads();
  }
  for (int o = 32; o >= 1; o >>= 1) {
    teps += __shfl_xor(teps, o, kWave);
    dag += __shfl_xor(dag, o, kWave);
  }
  if (ovf) atomicOr(&s_ovf, 1);
  if (lane == 0) {
    s_red[wave_id()][0] = teps;
    s_red[wave_id()][1] = dag;
  }
  __syncthreads();
  if (tid == 0) {
    u64 t = 0, g = 0;
    for (int w = 0; w < T / kWave; w++) {
      t += s_red[w][0];
      g += s_red[w][1];
    }
    a.part[3 * blockIdx.x] = t;
    a.part[3 * blockIdx.x + 1] = reached;
    a.part[3 * blockIdx.x + 2] = g;
    a.depth[blockIdx.x] = L;
    if (s_ovf) atomicOr(a.ovf, 1);
  }
}

// BC accumulation of one per-source launch: for every vertex, the rows'
// dependencies are summed in FP64 in source order, a run of sources with one
// target at a time, and each run's sum is added to its target once (the
// batched engine's lane-order reduction).
// With a cluster order nof (rows indexed by new ids), target entry v takes
// row entry nof[v].
__global__ __launch_bounds__(256) void k_ps_bc_reduce(float *const *rows, double *const *tgt, int nb,
                                                       int n, const int *nof) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < (size_t)n; v += stride) {
    const size_t x = nof ? (size_t)nof[v] : v;
    double acc = 0.0;
    for (int b = 0; b < nb; b++) {
      acc += (double)rows[b][x];
      double *t = tgt[b];
      if (b + 1 == nb || tgt[b + 1] != t) {
        if (t) t[v] += acc;
        acc = 0.0;
      }
    }
  }
}

// Rows computed in cluster order back to the caller's vertex order:
// dst[b][v] = src[b][nof[v]] (sequential writes, cluster-local reads).
// blockIdx.y selects the row.
template <class T>
__global__ __launch_bounds__(256) void k_ps_unpermute(T *const *src, T *const *dst, const int *nof, int n) {
  T *d = dst[blockIdx.y];
  if (!d) return;
  const T *s = src[blockIdx.y];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < (size_t)n; v += stride) d[v] = s[nof[v]];
}

}  // namespace gu
