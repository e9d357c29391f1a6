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

// Visit edge v->w at level L.  Returns true when this call claimed w.
template <bool SIGMA>
__device__ __forceinline__ bool ps_visit(int *d, u64 *sg, int w, int L, u64 sv, int &ovf, u64 &dag) {
  int dw = d[w];
  bool claimed = false;
  if (dw == INT_MAX) {
    dw = atomicCAS(&d[w], INT_MAX, L + 1);
    if (dw == INT_MAX) {
      claimed = true;
      dw = L + 1;
    }
  }
  if (SIGMA && dw == L + 1) {
    u64 old = atomicAdd(&sg[w], sv);
    ovf |= old + sv < sv;
    dag++;
  }
  return claimed;
}

// Expansion inside the workgroup: a thread walks each frontier vertex of
// degree <= kSerialDeg itself (linear_with_atomics.cu:120-157); larger
// adjacency lists are strip-mined by the whole wave, the lanes electing the
// highest lane that still has one (race_and_resolve, warp_based.cu:109-145).
template <bool SIGMA>
__global__ __launch_bounds__(1024) void k_persource(PsArgs a) {
  constexpr int kSerialDeg = 32;
  __shared__ unsigned s_q2n;
  __shared__ u64 s_red[16][2];
  u64 dag = 0;
  __shared__ int s_ovf;
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
  if (tid == 0) {
    s_q2n = 0;
    s_ovf = 0;
  }
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
    for (unsigned base = 0; base < qn; base += T) {
      const unsigned k = base + tid;
      const bool have = k < qn;
      const int v = have ? Q[k] : 0;
      int r = have ? a.R[v] : 0;
      const int re = have ? a.R[v + 1] : 0;
      const u64 sv = (SIGMA && have) ? ld_l2(&sg[v]) : 0ull;
      if (have) teps += (u64)(re - r);
      if (re - r <= kSerialDeg) {
        for (; r < re; r++) {
          const int w = a.C[r];
          if (ps_visit<SIGMA>(d, sg, w, L, sv, ovf, dag)) {
            unsigned pos = atomicAdd(&s_q2n, 1u);
            Q2[pos] = w;
          }
        }
        r = re;
      }
      // high-degree lists: the wave strip-mines the elected lane's list
      while (true) {
        const bool pend = r < re;
        if (!ballot(pend)) break;
        const int win = race_and_resolve_wave(pend);
        const int rw = readlane(r, win), rew = readlane(re, win);
        const u64 svw = SIGMA ? readlane64(sv, win) : 0ull;
        if (lane == win) r = re;
        for (int e0 = rw; e0 < rew; e0 += kWave) {
          const int ed = e0 + lane;
          bool claimed = false;
          int w = 0;
          if (ed < rew) {
            w = a.C[ed];
            claimed = ps_visit<SIGMA>(d, sg, w, L, svw, ovf, dag);
          }
          uint64_t m = ballot(claimed);
          if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            unsigned b = 0;
            if (lane == leader) b = atomicAdd(&s_q2n, (unsigned)__popcll(m));
            b = (unsigned)readlane((int)b, leader);
            if (claimed) Q2[b + rank_in(m)] = w;
          }
        }
      }
    }
    __syncthreads();
    const unsigned q2n = s_q2n;
    if (q2n == 0) break;
    reached += q2n;
    int *t = Q;
    Q = Q2;
    Q2 = t;
    qn = q2n;
    L++;
    __syncthreads();  // every thread has read s_q2n
    if (tid == 0) s_q2n = 0;
    __syncthreads();
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
  __shared__ unsigned s_q2n;
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
  if (tid == 0) {
    s_q2n = 0;
    s_ovf = 0;
  }
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
    for (unsigned base = qb; base < qe; base += T) {
      const unsigned k = base + tid;
      if (k < qe) {
        const int v = S[k];
        const int re = a.R[v + 1];
        const u64 sv = ld_l2(&sg[v]);
        teps += (u64)(re - a.R[v]);
        for (int r = a.R[v]; r < re; r++) {
          const int w = a.C[r];
          if (ps_visit<true>(d, sg, w, L, sv, ovf, dag)) {
            unsigned pos = atomicAdd(&s_q2n, 1u);
            S[qe + pos] = w;
          }
        }
      }
    }
    __syncthreads();
    const unsigned q2n = s_q2n;
    if (q2n == 0) break;
    reached += q2n;
    qb = qe;
    qe += q2n;
    L++;
    __syncthreads();
    if (tid == 0) {
      s_q2n = 0;
      E[L + 1] = (int)qe;
    }
    __syncthreads();
  }
  // backward, deepest level first (its vertices have no successors)
  for (int l = L - 1; l >= 1; l--) {
    const int b0 = E[l], b1 = E[l + 1];
    for (int i = b0 + tid; i < b1; i += T) {
      const int v = S[i];
      const float sv = (float)ld_l2(&sg[v]);
      float sum = 0.0f, comp = 0.0f;
      const int re = a.R[v + 1];
      for (int r = a.R[v]; r < re; r++) {
        const int w = a.C[r];
        if (ld_l2(&d[w]) == l + 1) {
          const float sw = (float)ld_l2(&sg[w]);
          ps_neumaier(sum, comp, (sv / sw) * (1.0f + dl[w]));
        }
      }
      dl[v] = sum + comp;
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
__global__ __launch_bounds__(256) void k_ps_bc_reduce(float *const *rows, double *const *tgt, int nb,
                                                       int n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < (size_t)n; v += stride) {
    double acc = 0.0;
    for (int b = 0; b < nb; b++) {
      acc += (double)rows[b][v];
      double *t = tgt[b];
      if (b + 1 == nb || tgt[b + 1] != t) {
        if (t) t[v] += acc;
        acc = 0.0;
      }
    }
  }
}

}  // namespace gu
