// gu_persource.hpp -- one workgroup per source, the whole level loop inside
// one launch.
//
// For high-diameter, bounded-degree graphs (grids, meshes, road-like graphs)
// the sources of a batch sit on disjoint frontier rings at every level, so
// sharing a level across sources buys nothing while thousands of levels each
// pay global passes.  Here a workgroup owns one source, as the reference's
// kernels do (one block per source, shuffle_based.cuh:104), and separates
// levels with workgroup barriers instead of kernel launches.  The frontier
// of a level lives in LDS (vertex and adjacency bounds, so the next level
// needs no row-offset loads), with entries beyond its capacity in an HBM
// queue per workgroup.
//
// Visit (count_shortest_paths' visit, shuffle_based.cu:130-141): a compare-
// and-swap claims an unvisited vertex for level L+1; every edge v->w with
// d[w] = L+1 then adds sigma[v] to sigma[w] atomically (exact, mod 2^64).
#pragma once
#include "gu_kernels.hpp"

namespace gu {

// How the workgroups of a per-source launch find their sources.
//  - One launch per batch (head == null): workgroup b takes source b
//    (nsrc = the grid); its rows are dist[b] / sigma[b].
//  - Queue (head != null): a grid of at most one workgroup per compute
//    unit takes sources from the device counter *head (cleared by the
//    host before the launch) until all nsrc are taken, so a call's sources
//    run as one grid and no workgroup waits for another's deeper search
//    (the reference's block-strided source loop, shuffle_based.cuh:104).
//    A null entry of the row tables means the workgroup's scratch row
//    (dscr / sscr + slot * n).  With nof set, the rows were searched in
//    cluster order and the workgroup writes them back itself when its
//    search ends: ddst[j][v] = d[nof[v]] (INT_MAX where not reached),
//    sdst[j][v] = sigma[nof[v]] (0 where not reached); a null table or
//    entry is not written.
struct PsQueue {
  int nsrc;
  int *head;
  int *dscr;          // [grid][n] distance scratch rows (int, or 16-bit)
  u64 *sscr;          // [grid][n] path-count scratch rows
  const int *nof;     // input id -> cluster id, or null (no write-back)
  int *const *ddst;   // [nsrc] the caller's distance rows
  u64 *const *sdst;   // [nsrc] the caller's path-count rows
};

// The next source of workgroup `slot` (uniform across the workgroup), or
// nsrc when none is left.
__device__ __forceinline__ int ps_next_job(const PsQueue &q, int *s_job, bool first) {
  if (!q.head) return first ? (int)blockIdx.x : q.nsrc;
  if (threadIdx.x == 0) *s_job = atomicAdd(q.head, 1);
  __syncthreads();
  return *s_job;
}

// Row j's distances and path counts back to the caller's vertex order (see
// PsQueue), by the workgroup that searched them; `none` marks an
// unreached vertex in the distance row.  Called after a workgroup
// barrier, so the rows' values written by every wave are visible.
template <class DT>
__device__ __forceinline__ void ps_writeback(const PsQueue &q, int j, gp<const DT> ds, gp<const u64> ss, int n,
                                             DT none) {
  gp<int> dd = glob(q.ddst ? q.ddst[j] : nullptr);
  gp<u64> sd = glob(q.sdst && ss ? q.sdst[j] : nullptr);
  gp<const int> nof = glob(q.nof);
  if (!dd && !sd) return;
  // One workgroup moves a whole row pair, so the steps (U vertices a
  // thread) are pipelined: step i+2's map entries and step i+1's row
  // gathers are issued before step i's stores.  One in-order counter covers
  // loads and stores on this hardware, so a load issued after a store is
  // waited for only with that store; in this order no wait covers a store.
  // Every load is issued (ids past the row read entry n-1), so the loop has
  // no branch around its loads.
  constexpr int U = 8;
  const int T = blockDim.x, tid = threadIdx.x;
  const int steps = (n + U * T - 1) / (U * T);
  auto at = [&](int i, int u) { return i * U * T + u * T + tid; };
  auto cl = [&](int v) { return v < n ? v : n - 1; };
  int xa[U];
  DT da[U];
  u64 sa[U];
  {
    int x0[U];
#pragma unroll
    for (int u = 0; u < U; u++) x0[u] = nof[cl(at(0, u))];
#pragma unroll
    for (int u = 0; u < U; u++) xa[u] = nof[cl(at(1, u))];
#pragma unroll
    for (int u = 0; u < U; u++) {
      da[u] = ds[x0[u]];
      sa[u] = sd ? ss[x0[u]] : 0ull;
    }
  }
  for (int i = 0; i < steps; i++) {
    int xb[U];
    DT db[U];
    u64 sb[U];
#pragma unroll
    for (int u = 0; u < U; u++) xb[u] = nof[cl(at(i + 2, u))];
#pragma unroll
    for (int u = 0; u < U; u++) {
      db[u] = ds[xa[u]];
      sb[u] = sd ? ss[xa[u]] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int v = at(i, u);
      if (v >= n) continue;
      const bool r = da[u] != none;
      if (dd) dd[v] = r ? (int)da[u] : INT_MAX;
      if (sd) sd[v] = r ? sa[u] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      xa[u] = xb[u];
      da[u] = db[u];
      sa[u] = sb[u];
    }
  }
}

struct PsArgs {
  const int *R;
  const int *C;
  const int4 *LE;    // degree <= 4 graphs: neighbours per vertex, -1 padded
  int n;
  const int *srcs;   // [nsrc]
  int *const *dist;  // [nsrc] distance rows of n entries
  u64 *const *sigma; // [nsrc] path-count rows, or null
  int *queues;       // [nsrc][2n+2]: frontier entries beyond the LDS capacity
  int *depth;        // [nsrc] deepest level reached
  u64 *part;         // [nsrc][4] (TEPS units, reached vertices, DAG edges,
                     //   1 if one of the source's path counts wrapped)
  int *ovf;          // set when a path count wrapped
  PsQueue q;         // source queue and in-kernel write-back
};

// Relaxed atomic loads at workgroup scope for the distance and path-count
// rows that other waves of the workgroup claim (compare-and-swap) and store
// during the search.  One workgroup owns a source's rows, and every value a
// level relies on was written before a workgroup barrier, so workgroup
// scope suffices; the loads may then hit in the CU's L1.  Against agent
// scope this took a 1-source search with path counts on a 2048^2 grid from
// 46.0 to 37.5 ms and 1-source Brandes on a 1024^2 grid from 16.2 to
// 13.8 ms, with the same HBM traffic.
__device__ __forceinline__ u64 ld_wg(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int ld_wg(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ u64 ld_wg(gp<const u64> p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int ld_wg(gp<const int> p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

#ifndef GU_PS_CAP
#define GU_PS_CAP 3072
#endif

// LDS frontier of the per-source kernels: two buffers of kPsCapLds entries
// (vertex, first and one-past-last adjacency index); a level's entries
// beyond that go to `spill` (vertices only) in HBM.
constexpr int kPsCapLds = GU_PS_CAP;
struct PsFront {
  int v[2][kPsCapLds];
  int r[2][kPsCapLds];
  int e[2][kPsCapLds];
};

// Vertices of degree above this are strip-mined by a whole wave.
constexpr int kPsSerialDeg = 32;

// Reads entry k of buffer b: (vertex, adjacency bounds).
__device__ __forceinline__ void ps_entry(const PsFront &f, gp<const int> spill, const int *R, int b,
                                         unsigned k, int &v, int &r, int &re) {
  if (k < (unsigned)kPsCapLds) {
    v = f.v[b][k];
    r = f.r[b][k];
    re = f.e[b][k];
  } else {
    v = spill[k - kPsCapLds];
    r = R[v];
    re = R[v + 1];
  }
}

// Appends claimed vertex w at position p of the next buffer.  The bounds
// are loaded here, while the level's path-count adds are in flight.
__device__ __forceinline__ void ps_push(PsFront &f, gp<int> spill, gp<int> S, const int *R, int b, unsigned p,
                                        int w) {
  if (S) S[p] = w;
  if (p < (unsigned)kPsCapLds) {
    f.v[b][p] = w;
    f.r[b][p] = R[w];
    f.e[b][p] = R[w + 1];
  } else {
    spill[p - kPsCapLds] = w;
  }
}

// Distances of the neighbours w[0..B) (o = -1 for padding lanes, w = -1).
// Every lane loads and the value is selected afterwards: with the loads
// under branches the wait-count pass serialised them (a wait after each
// conditional load; branch-free, 64-source diameter sampling on a 2048^2
// grid went from 39.2 to 33.8 ms).  Padding lanes read entry 0 of the row,
// one line for the whole wave (re-reading a line the wave requests anyway
// measured 4% slower with 256 sources in flight).
template <bool AT, int B>
__device__ __forceinline__ void ps_load_d(gp<const int> d, const int *w, int *o) {
#pragma unroll
  for (int j = 0; j < B; j++) {
    gp<const int> p = d + (w[j] >= 0 ? w[j] : 0);
    const int t = AT ? ld_wg(p) : *p;
    o[j] = w[j] >= 0 ? t : -1;
  }
}

// Path counts of the parents among the neighbours (distance L-1 at the
// first read: a claim only turns INT_MAX into L+1, so the parents are known
// before the claims and these loads issue beside them), branch-free as in
// ps_load_d (the other lanes read entry 0); x = 0 off the parents.
template <int B>
__device__ __forceinline__ void ps_load_parents(gp<const u64> sg, const int *w, const int *o, int L, bool *par,
                                                u64 *x) {
#pragma unroll
  for (int j = 0; j < B; j++) par[j] = w[j] >= 0 && L >= 1 && o[j] == L - 1;
#pragma unroll
  for (int j = 0; j < B; j++) {
    const u64 t = ld_wg(sg + (par[j] ? w[j] : 0));
    x[j] = par[j] ? t : 0ull;
  }
}

// Visits of frontier vertex v (level L) over its adjacency [r, re) in
// steps of 8: the column loads, then reads of d, then compare-and-swap
// claims INT_MAX -> L+1 of the unvisited neighbours, whose positions one LDS
// atomic per step reserves (S, for Brandes, also records them in HBM).
// Path counts (count_shortest_paths' visit, shuffle_based.cu:130-141: every
// edge u->w with d[w] = d[u]+1 adds sigma[u] to sigma[w]) are formed in one
// of two ways:
//  - PULL (undirected graphs, where the in-neighbours are the adjacency):
//    sigma[v] is the sum of sigma[u] over the neighbours u with d[u] = L-1,
//    read in the same pass, so no atomics; d and sigma are read with
//    workgroup-scope atomic loads (ld_wg);
//  - push (directed graphs): each edge v->w with d[w] = L+1 adds sigma[v]
//    to sigma[w] atomically.
// Both sums are exact mod 2^64; `ovf` flags a wrap.
// Lanes past the adjacency carry w = -1 and o = -1; every use of o that a
// level can equal (o == L-1 is -1 at level 0) is guarded by w >= 0.
template <bool SIGMA, bool PULL>
__device__ __forceinline__ void ps_visit(const int *R, const int *C, gp<int> d, gp<u64> sg, int v, int r, int re,
                                         int L, u64 sv, int &ovf, u64 &dag, unsigned *cnt, PsFront &f,
                                         gp<int> spill, gp<int> S, int nbuf) {
  constexpr int B = 8;
  u64 acc = 0;
  for (; r < re; r += B) {
    int w[B], o[B];
#pragma unroll
    for (int j = 0; j < B; j++) w[j] = r + j < re ? C[r + j] : -1;
    ps_load_d<SIGMA && PULL, B>(d, w, o);
    bool par[B];
    u64 x[B];
    if (SIGMA && PULL)
      ps_load_parents<B>(sg, w, o, L, par, x);
    else
      for (int j = 0; j < B; j++) par[j] = false, x[j] = 0;
#pragma unroll
    for (int j = 0; j < B; j++)
      if (o[j] == INT_MAX) o[j] = cas_g(d + w[j], INT_MAX, L + 1);  // INT_MAX back: claimed
    unsigned claims = 0;
#pragma unroll
    for (int j = 0; j < B; j++) claims += o[j] == INT_MAX;
    unsigned pos = claims ? atomicAdd(cnt, claims) : 0u;
    if (SIGMA && !PULL) {
#pragma unroll
      for (int j = 0; j < B; j++) x[j] = (o[j] == INT_MAX || o[j] == L + 1) ? add_g(sg + w[j], sv) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < B; j++)
      if (o[j] == INT_MAX) ps_push(f, spill, S, R, nbuf, pos++, w[j]);
    if (SIGMA) {
#pragma unroll
      for (int j = 0; j < B; j++) {
        if (PULL && par[j]) {
          acc += x[j];
          ovf |= acc < x[j];
          dag++;
        }
        if (!PULL && (o[j] == INT_MAX || o[j] == L + 1)) {
          ovf |= x[j] + sv < sv;
          dag++;
        }
      }
    }
  }
  if (SIGMA && PULL && v >= 0 && L >= 1) sg[v] = acc;
}

// ps_visit on a graph of degree at most 4 kept as one int4 of neighbours
// per vertex (-1 padded): the visit is one 16-byte load, and a claimed
// vertex enters the next frontier as its id alone, with no row-offset loads
// before the level's barrier.
template <bool SIGMA, bool PULL, int NV>
__device__ __forceinline__ void ps_visit_ell(const int4 *LE, gp<int> d, gp<u64> sg, const int *vv, int L,
                                             const u64 *sv, int &ovf, u64 &dag, u64 &teps, unsigned *cnt,
                                             PsFront &f, gp<int> spill, gp<int> S, int nbuf) {
  constexpr int B = 4 * NV;
  int w[B], o[B];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const int4 q = vv[i] >= 0 ? LE[vv[i]] : make_int4(-1, -1, -1, -1);
    w[4 * i] = q.x;
    w[4 * i + 1] = q.y;
    w[4 * i + 2] = q.z;
    w[4 * i + 3] = q.w;
  }
#pragma unroll
  for (int j = 0; j < B; j++) teps += w[j] >= 0;
  ps_load_d<SIGMA && PULL, B>(d, w, o);
  // (Loading every neighbour's path count beside its distance, to use it
  // only for parents, saves a round trip but measured 6% slower on a grid:
  // the extra lines cost more.)
  bool par[B];
  u64 x[B];
  if (SIGMA && PULL)
    ps_load_parents<B>(sg, w, o, L, par, x);
  else
    for (int j = 0; j < B; j++) par[j] = false, x[j] = 0;
#pragma unroll
  for (int j = 0; j < B; j++)
    if (o[j] == INT_MAX) o[j] = cas_g(d + w[j], INT_MAX, L + 1);  // INT_MAX back: claimed
  unsigned claims = 0;
#pragma unroll
  for (int j = 0; j < B; j++) claims += o[j] == INT_MAX;
  unsigned pos = wave_reserve(cnt, claims);
  if (SIGMA && !PULL) {
#pragma unroll
    for (int j = 0; j < B; j++)
      x[j] = (o[j] == INT_MAX || o[j] == L + 1) ? add_g(sg + w[j], sv[j / 4]) : 0ull;
  }
#pragma unroll
  for (int j = 0; j < B; j++) {
    if (o[j] != INT_MAX) continue;
    if (S) S[pos] = w[j];
    if (pos < (unsigned)kPsCapLds)
      f.v[nbuf][pos] = w[j];
    else
      spill[pos - kPsCapLds] = w[j];
    pos++;
  }
  if (SIGMA) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      u64 acc = 0;
#pragma unroll
      for (int j = 4 * i; j < 4 * i + 4; j++) {
        if (PULL && par[j]) {
          acc += x[j];
          ovf |= acc < x[j];
          dag++;
        }
        if (!PULL && (o[j] == INT_MAX || o[j] == L + 1)) {
          ovf |= x[j] + sv[i] < sv[i];
          dag++;
        }
      }
      if (PULL && L >= 1 && vv[i] >= 0) sg[vv[i]] = acc;
    }
  }
}

// A wave strip-mines the adjacency lists longer than kPsSerialDeg of its
// lanes, electing the highest lane that still has one (race_and_resolve,
// warp_based.cu:109-145); in PULL mode the lanes' partial path counts are
// summed across the wave and the elected lane stores them.
template <bool SIGMA, bool PULL>
__device__ __forceinline__ void ps_visit_wave(const int *R, const int *C, gp<int> d, gp<u64> sg, int v, int rr,
                                              int rre, int L, u64 sv, int &ovf, u64 &dag, unsigned *cnt,
                                              PsFront &f, gp<int> spill, gp<int> S, int nbuf) {
  const int lane = lane_id();
  while (true) {
    const bool pend = rr < rre;
    if (!ballot(pend)) break;
    const int win = race_and_resolve_wave(pend);
    const int rw = readlane(rr, win), rew = readlane(rre, win), vw = readlane(v, win);
    const u64 svw = (SIGMA && !PULL) ? readlane64(sv, win) : 0ull;
    if (lane == win) rr = rre;
    u64 acc = 0;
    int wrap = 0;
    for (int e0 = rw; e0 < rew; e0 += kWave) {
      const int ed = e0 + lane;
      int w = -1, o = -1;
      if (ed < rew) {
        w = C[ed];
        o = (SIGMA && PULL) ? ld_wg(d + w) : d[w];
        if (o == INT_MAX) o = cas_g(d + w, INT_MAX, L + 1);
      }
      const bool claimed = o == INT_MAX;
      uint64_t m = ballot(claimed);
      unsigned b = 0;
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        if (lane == leader) b = atomicAdd(cnt, (unsigned)__popcll(m));
        b = (unsigned)readlane((int)b, leader);
      }
      if (SIGMA && PULL && w >= 0 && L >= 1 && o == L - 1) {
        const u64 x = ld_wg(sg + w);
        acc += x;
        wrap |= acc < x;
        dag++;
      }
      if (SIGMA && !PULL && (claimed || o == L + 1)) {
        const u64 old = add_g(sg + w, svw);
        ovf |= old + svw < svw;
        dag++;
      }
      if (claimed) ps_push(f, spill, S, R, nbuf, b + rank_in(m), w);
    }
    if (SIGMA && PULL) {
      for (int o = 32; o >= 1; o >>= 1) {
        const u64 y = __shfl_xor(acc, o, kWave);
        acc += y;
        wrap |= acc < y;
      }
      ovf |= wrap;
      if (lane == win && L >= 1) sg[vw] = acc;
    }
  }
}

// The level loop of the search and Brandes kernels.  Levels are separated
// by one workgroup barrier: the level counters rotate over three LDS words,
// the one two levels back being cleared for the next level.  With BC every
// level's vertices are also appended to S in HBM (E[l] = start of level l),
// and a frontier's entries beyond the LDS capacity are read back from S;
// otherwise they go to the two `spill` queues.  Returns the deepest level.
template <bool SIGMA, bool PULL, bool BC, bool ELL = false>
__device__ __forceinline__ int ps_levels(const int *R, const int *C, int n, int s, gp<int> d, gp<u64> sg, gp<int> spill0,
                         gp<int> S, gp<int> E, PsFront &f, unsigned *s_cnt, int &ovf, u64 &teps, u64 &reached,
                         u64 &dag, const int4 *LE = nullptr) {
  const int tid = threadIdx.x, T = blockDim.x;
  if (tid < 3) s_cnt[tid] = 0;
  if (tid == 0) {
    d[s] = 0;
    if (SIGMA) sg[s] = 1;
    f.v[0][0] = s;
    f.r[0][0] = R[s];
    f.e[0][0] = R[s + 1];
    if (BC) {
      S[0] = s;
      E[0] = 0;
      E[1] = 1;
    }
  }
  __syncthreads();
  unsigned qn = 1, qb = 0, qs = 1;  // frontier size; its start and end in S
  int L = 0, b = 0;
  reached = 1;
  while (true) {
    unsigned *cnt = &s_cnt[L % 3];
    if (tid == 0) s_cnt[(L + 1) % 3] = 0;
    gp<int> Snext = BC ? S + qs : nullptr;
    gp<int> sp_cur = BC ? S + qb + kPsCapLds : spill0 + (size_t)b * n;
    gp<int> sp_next = BC ? S + qs + kPsCapLds : spill0 + (size_t)(b ^ 1) * n;
    if (ELL) {  // one frontier vertex per thread and slice (2 or 3, with
                // their loads side by side, measured 8-9% slower on a grid)
      constexpr int NV = 1;
      for (unsigned base = 0; base < qn; base += NV * T) {
        int vv[NV];
        u64 sv[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) {
          const unsigned k = base + i * T + tid;
          vv[i] = k >= qn ? -1 : (k < (unsigned)kPsCapLds ? f.v[b][k] : sp_cur[k - kPsCapLds]);
          sv[i] = (SIGMA && !PULL && vv[i] >= 0) ? ld_wg(sg + vv[i]) : 0ull;
        }
        ps_visit_ell<SIGMA, PULL, NV>(LE, d, sg, vv, L, sv, ovf, dag, teps, cnt, f, sp_next, Snext, b ^ 1);
      }
    }
    for (unsigned base = 0; !ELL && base < qn; base += T) {
      const unsigned k = base + tid;
      int v = -1, r = 0, re = 0;
      if (k < qn) ps_entry(f, sp_cur, R, b, k, v, r, re);
      const u64 sv = (SIGMA && !PULL && v >= 0) ? ld_wg(sg + v) : 0ull;
      teps += (u64)(re - r);
      const bool big = re - r > kPsSerialDeg;
      ps_visit<SIGMA, PULL>(R, C, d, sg, v, big ? 0 : r, big ? 0 : re, L, sv, ovf, dag, cnt, f, sp_next,
                            Snext, b ^ 1);
      ps_visit_wave<SIGMA, PULL>(R, C, d, sg, v, big ? r : 0, big ? re : 0, L, sv, ovf, dag, cnt, f, sp_next,
                                 Snext, b ^ 1);
    }
    __syncthreads();
    const unsigned q2n = *cnt;
    if (q2n == 0) break;
    reached += q2n;
    qn = q2n;
    qb = qs;
    qs += q2n;
    b ^= 1;
    L++;
    if (BC && tid == 0) E[L + 1] = (int)qs;
  }
  return L;
}

// Workgroup sums of the per-thread counters; thread 0 writes entry j (the
// source's index in the call) of part / depth (part[3]: a path count of
// this source wrapped).
__device__ __forceinline__ void ps_finish(u64 teps, u64 dag, u64 reached, int L, int ovf, int *s_ovf,
                                          u64 (*s_red)[2], u64 *part, int *depth, int *ovf_out,
                                          int j = blockIdx.x) {
  const int tid = threadIdx.x, T = blockDim.x;
  for (int o = 32; o >= 1; o >>= 1) {
    teps += __shfl_xor(teps, o, kWave);
    dag += __shfl_xor(dag, o, kWave);
  }
  if (ovf) atomicOr(s_ovf, 1);
  if (lane_id() == 0) {
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
    part[4 * (size_t)j] = t;
    part[4 * (size_t)j + 1] = reached;
    part[4 * (size_t)j + 2] = g;
    part[4 * (size_t)j + 3] = *s_ovf ? 1ull : 0ull;
    depth[j] = L;
    if (*s_ovf) atomicOr(ovf_out, 1);
  }
}

template <bool SIGMA, bool PULL, bool ELL = false>
__global__ __launch_bounds__(1024) void k_persource(PsArgs a) {
  __shared__ PsFront f;
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf, s_job;
  const int tid = threadIdx.x, T = blockDim.x;
  const size_t slot = blockIdx.x, n = (size_t)a.n;
  gp<int> spill = glob(a.queues + slot * (2 * n + 2));
  for (int j = ps_next_job(a.q, &s_job, true); j < a.q.nsrc; j = ps_next_job(a.q, &s_job, false)) {
    const int s = a.srcs[j];
    gp<int> d = glob(a.dist && a.dist[j] ? a.dist[j] : a.q.dscr + slot * n);
    gp<u64> sg = glob(SIGMA ? (a.sigma[j] ? a.sigma[j] : a.q.sscr + slot * n) : nullptr);
    for (int i = tid; i < a.n; i += T) {
      d[i] = INT_MAX;
      if (SIGMA) sg[i] = 0;
    }
    if (tid == 0) s_ovf = 0;
    __syncthreads();
    int ovf = 0;
    u64 teps = 0, reached = 1, dag = 0;
    const int L = ps_levels<SIGMA, PULL, false, ELL>(a.R, a.C, a.n, s, d, sg, spill, nullptr, nullptr, f, s_cnt,
                                                     ovf, teps, reached, dag, a.LE);
    ps_finish(teps, dag, reached, L, ovf, &s_ovf, s_red, a.part, a.depth, a.ovf, j);
    if (a.q.nof) ps_writeback<int>(a.q, j, d, sg, a.n, INT_MAX);
    __syncthreads();  // rows and s_job read before the next source
  }
}

// Brandes on one workgroup per source (the dependency accumulation of
// bc.cu:11-57 in pull form, as the reference's GPU backward,
// shuffle_based.cu:217-441): the forward pass above with the BFS order kept
// as a stack S[] of levels (E[l] = start of level l in S), then the levels
// from the deepest down, each vertex v summing over successors w (d[w] =
// d[v]+1, adjacency order) (sigma_v / sigma_w)(1 + delta_w) in FP32 with a
// Neumaier compensated sum; delta of the source stays 0 (bc.cu:50).
// The backward reads distances and path counts with ld_wg, as the forward.
struct PsBcArgs {
  const int *R;
  const int *C;
  const int4 *LE;       // degree <= 4 graphs: the forward's neighbour table
  int n;
  const int *srcs;
  int *const *dist;     // [nsrc] scratch rows
  u64 *const *sigma;    // [nsrc] scratch rows
  float *const *delta;  // [nsrc] dependency rows
  float *const *quot;   // [nsrc] scratch rows: (1 + delta) / sigma of each
                        //   vertex once its delta is final
  int *stack;           // [nsrc][2n+2]: S, then level starts E (HBM)
  int *depth;
  u64 *part;
  int *ovf;
  int lds_flags;        // kLds* graph facts for the level-set kernels
  const u64 *LE8;       // or null: the neighbour table as int16 deltas
  u64 *slog;            // level-set kernels: [grid][2 * kLdsLog] path-count logs
};

__device__ __forceinline__ void ps_neumaier(float &sum, float &comp, float x) {
  float t = sum + x;
  if (fabsf(sum) >= fabsf(x))
    comp += (sum - t) + x;
  else
    comp += (x - t) + sum;
  sum = t;
}

// Successors among the neighbours w[0..B) of a level-l vertex (d = l+1,
// read with ld_wg) and their q, branch-free as in ps_load_d: every lane
// loads (the others read entry 0) and the values are selected afterwards;
// qw = 0 off the successors.
template <int B>
__device__ __forceinline__ void ps_successors(gp<const int> d, gp<const float> ql, const int *w, int l, bool *on,
                                              float *qw) {
  int o[B];
  ps_load_d<true, B>(d, w, o);
#pragma unroll
  for (int j = 0; j < B; j++) on[j] = w[j] >= 0 && o[j] == l + 1;
#pragma unroll
  for (int j = 0; j < B; j++) {
    const float t = ql[on[j] ? w[j] : 0];
    qw[j] = on[j] ? t : 0.0f;
  }
}

template <bool PULL, bool ELL = false>
__global__ __launch_bounds__(1024) void k_ps_bc(PsBcArgs a) {
  __shared__ PsFront f;
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf;
  const int tid = threadIdx.x, T = blockDim.x;
  const int s = a.srcs[blockIdx.x];
  gp<int> d = glob(a.dist[blockIdx.x]);
  gp<u64> sg = glob(a.sigma[blockIdx.x]);
  gp<float> dl = glob(a.delta[blockIdx.x]);
  gp<float> ql = glob(a.quot[blockIdx.x]);
  gp<int> S = glob(a.stack + (size_t)blockIdx.x * (2 * (size_t)a.n + 2));
  gp<int> E = S + a.n;
  // frontier entries beyond the LDS capacity are read back from S itself:
  // a level's vertices occupy S[E[l], E[l+1]) in frontier order
  for (int i = tid; i < a.n; i += T) {
    d[i] = INT_MAX;
    sg[i] = 0;
    dl[i] = 0.0f;
  }
  if (tid == 0) s_ovf = 0;
  __syncthreads();
  int ovf = 0;
  u64 teps = 0, reached = 1, dag = 0;
  const int L = ps_levels<true, PULL, true, ELL>(a.R, a.C, a.n, s, d, sg, nullptr, S, E, f, s_cnt, ovf, teps,
                                                 reached, dag, a.LE);
  __syncthreads();
  // backward, deepest level first: its vertices have no successors (delta
  // 0, q = 1 / sigma); then each thread sums the successors of one vertex in
  // adjacency order, sigma_v * q_w (q_w = (1 + delta_w) / sigma_w, stored
  // when delta_w became final), loading a chunk's distances together, then
  // its successors' q together
  for (int i = E[L] + tid; i >= 0 && L >= 1 && i < E[L + 1]; i += T) {
    const int v = S[i];
    ql[v] = (1.0f + 0.0f) / (float)ld_wg(sg + v);
  }
  __syncthreads();
  for (int l = L - 1; l >= 1; l--) {
    const int b0 = E[l], b1 = E[l + 1];
    for (int i = b0 + tid; i < b1; i += T) {
      const int v = S[i];
      const float sv = (float)ld_wg(sg + v);
      float sum = 0.0f, comp = 0.0f;
      if (ELL) {  // the neighbours in one load, in adjacency order
        const int4 q4 = a.LE[v];
        const int w[4] = {q4.x, q4.y, q4.z, q4.w};
        bool on[4];
        float qw[4];
        ps_successors<4>(d, ql, w, l, on, qw);
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (on[j]) ps_neumaier(sum, comp, sv * qw[j]);
        const float dv = sum + comp;
        dl[v] = dv;
        ql[v] = (1.0f + dv) / sv;
        continue;
      }
      int r = a.R[v];
      const int re = a.R[v + 1];
      constexpr int B = 8;
      for (; r < re; r += B) {
        int w[B];
        bool on[B];
        float qw[B];
#pragma unroll
        for (int j = 0; j < B; j++) w[j] = r + j < re ? a.C[r + j] : -1;
        ps_successors<B>(d, ql, w, l, on, qw);
#pragma unroll
        for (int j = 0; j < B; j++)
          if (on[j]) ps_neumaier(sum, comp, sv * qw[j]);
      }
      const float dv = sum + comp;
      dl[v] = dv;
      ql[v] = (1.0f + dv) / sv;
    }
    __syncthreads();
  }
  ps_finish(teps, dag, reached, L, ovf, &s_ovf, s_red, a.part, a.depth, a.ovf);
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
      acc += (double)glob(rows[b])[x];
      gp<double> t = glob(tgt[b]);
      if (b + 1 == nb || glob(tgt[b + 1]) != t) {
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
  gp<T> d = glob(dst[blockIdx.y]);
  if (!d) return;
  gp<const T> s = glob((const T *)src[blockIdx.y]);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < (size_t)n; v += stride) d[v] = s[nof[v]];
}

// ---- undirected degree <= 4 graphs: the last levels as LDS hash sets ----
//
// On an undirected graph every neighbour of a vertex at BFS level L lies at
// level L-1, L or L+1 (or is not reached yet), so a source's search needs
// only the vertex sets of the three levels around the current one, not its
// distance row.  Those sets live in the workgroup's LDS as four rotating
// open-addressing hash sets (4-slot buckets probed linearly, vertex ids, -1
// empty): level L
// reads the sets of L-1 (parents) and L (same level), inserts claims into
// the set of L+1 with an LDS compare-and-swap (each vertex claimed once),
// and clears the set that held L-2, which takes the claims of L+2 after
// this level's barrier.  Per frontier vertex the only global reads are its
// int4 of neighbours and its parents' path counts; distances and counts
// are stored once, when a vertex is expanded.  The parents' counts are read
// from the previous level's log: a vertex's count goes both to its row and
// to entry k of its level's log (k: its index in the level's frontier),
// and a level set keeps each vertex's k beside its key, so the reads of a
// level fall in one compact block written one level before (8 bytes per
// vertex of that level) rather than across the row.  A level set that
// cannot place a claim within kLdsProbe buckets stops the source (flag 2
// in part[3]) and the host reruns the launch on the distance-reading
// kernels, so results never depend on LDS capacity.
constexpr int kLdsBucketBits = 11;
constexpr int kLdsBuckets = 1 << kLdsBucketBits;  // 4-slot buckets per level set (8192 slots)
constexpr int kLdsProbe = 15;                     // buckets probed before giving up
constexpr int kLdsFront = 3072;                   // LDS frontier entries per buffer
constexpr int kLdsSlots = 3;                      // frontier vertices a thread loads at once
// Slots.  Searches without path counts keep vertex ids in the slots (home
// bucket: the top 11 bits of w * kLdsMul mod 2^32).  With path counts the
// sets are keyed and hold vertex ids below 2^23: h = w * kLdsKeyMul mod
// 2^23 is a bijection on them (odd multiplier); both multipliers are the
// word size over the golden ratio, so consecutive ids spread evenly over
// the buckets.  h's top 11 bits are w's home bucket,
// and its low 12 bits r stay in the slot with the probe distance p (the
// bucket's distance from home): key = r | p << 12.  Two ids with one key in
// one bucket have the same h, so they are the same id.  A keyed slot holds
// key | k << 16 (k < 8192, the vertex's index in its level's frontier, set
// by the claim's winner once known), a value below 2^29; -1 marks an empty
// slot.  With p < 15 a key stays below 0xF000, so a slot matches a key
// when its low 16 bits equal it (one 16-bit compare): an empty slot's low
// half, 0xFFFF, is no key.
constexpr int kLdsIdBits = 23;
constexpr int kLdsMaxN = 1 << kLdsIdBits;  // graphs the keyed level sets can hold
constexpr int kLdsRBits = kLdsIdBits - kLdsBucketBits;
constexpr unsigned kLdsMul = 0x9E3779B1u;     // 2^32 / golden ratio, odd
constexpr unsigned kLdsKeyMul = 0x4F1BBDu;    // 2^23 / golden ratio, odd
constexpr int kLdsLog = 4 * kLdsBuckets;  // path-count log entries per level (one per slot)
// Facts about the graph that the host measured when it built the handle.
// Bipartite: BFS levels are independent sets, so a neighbour of a level-L
// vertex is never at level L and the same-level lookup is skipped.
// Connected: every vertex is reached, so the distance row needs no fill.
// 16-bit distances of the level-set search's scratch rows.
constexpr unsigned short kD16None = 0xffff;
constexpr int kD16Max = 0xfffe;
// With path counts the level-set search keeps one 16-byte record per
// vertex instead of a distance row and a path-count row: (path count low
// and high words, distance, unused), distance -1 where not reached.  One
// store writes a vertex's pair into one 32-byte sector; a 2-byte distance
// store to a row of its own dirtied a second sector per vertex (2.5 ms of
// a 256-source 2048^2 pass).
// (A clang vector type: it can be stored through a global-address-space
// pointer, which HIP's int4 class cannot.)
typedef int ps_rec_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ ps_rec_t ps_rec(u64 sigma, int d) {
  return ps_rec_t{(int)(unsigned)sigma, (int)(unsigned)(sigma >> 32), d, 0};
}

// Compact neighbour table: four int16 deltas w - v in one u64 (lowest 16
// bits first), kLe8None for an empty slot.
constexpr int kLe8None = -32768;
__device__ __forceinline__ int le8_one(u64 p, int j, int v) {
  const int dl = (int)(short)(unsigned short)(p >> (16 * j));
  return dl == kLe8None ? -1 : v + dl;
}
__device__ __forceinline__ int4 le8_decode(u64 p, int v) {
  return make_int4(le8_one(p, 0, v), le8_one(p, 1, v), le8_one(p, 2, v), le8_one(p, 3, v));
}
// The neighbours of x (x >= 0) from whichever table the graph has.
__device__ __forceinline__ int4 le_load(const int4 *LE, const u64 *LE8, int x) {
  return LE8 ? le8_decode(LE8[x], x) : LE[x];
}
// The same in two steps, so that a load issued ahead is not waited for
// until its value is used: the table entry of x (entry 0 for x < 0, which
// decodes to no neighbours), then its neighbours.  The table is a template
// argument (E8: the int16-delta table), so no run-time branch between the
// two loads sits in front of the wait-count pass.
template <bool E8>
__device__ __forceinline__ int4 le_fetch(const int4 *LE, const u64 *LE8, int x) {
  const int y = x >= 0 ? x : 0;
  if (E8) {
    const int2 t = reinterpret_cast<const int2 *>(LE8)[y];
    return make_int4(t.x, t.y, 0, 0);
  }
  return LE[y];
}
template <bool E8>
__device__ __forceinline__ int4 le_decode(int x, const int4 r) {
  if (x < 0) return make_int4(-1, -1, -1, -1);
  if (E8) return le8_decode(((u64)(unsigned)r.y << 32) | (unsigned)r.x, x);
  return r;
}
constexpr int kLdsBipartite = 1;
constexpr int kLdsConnected = 2;
struct PsLds {
  int4 tab[4][kLdsBuckets];  // 128 KB
  int fv[2][kLdsFront];      // 24 KB; entries beyond go to HBM
};

__device__ __forceinline__ unsigned lds_bucket(int w) { return ((unsigned)w * kLdsMul) >> (32 - kLdsBucketBits); }
__device__ __forceinline__ bool q4_has(const int4 q, int w) { return q.x == w || q.y == w || q.z == w || q.w == w; }
__device__ __forceinline__ unsigned lds_hash(int w) { return ((unsigned)w * kLdsKeyMul) & ((1u << kLdsIdBits) - 1); }
__device__ __forceinline__ bool lds_match(int x, unsigned key) { return (unsigned short)x == (unsigned short)key; }
__device__ __forceinline__ bool q4_open(const int4 q) { return q.x < 0 || q.y < 0 || q.z < 0 || q.w < 0; }

// Buckets are probed linearly, one 16-byte LDS read each, and fill from
// their first slot: a key moves on to the next bucket only when its bucket
// is full, and nothing leaves a set during a level, so a bucket with a free
// slot ends the search.  At most kLdsProbe buckets (an insert gives up
// beyond).  With levels of up to 4096 vertices (load <= 1/2) almost every
// lookup reads one bucket, so the lanes of a wave stay in step.  Returns
// w's frontier index (KEYED; else 0) when w is in the set, else -1.
template <bool KEYED>
__device__ __forceinline__ int lds_find(const int4 *t, int w) {
  if (!KEYED) {
    unsigned b = lds_bucket(w);
    for (int p = 0; p < kLdsProbe; p++) {
      const int4 q = t[b];
      if (q4_has(q, w)) return 0;
      if (q4_open(q)) return -1;
      b = (b + 1) & (kLdsBuckets - 1);
    }
    return -1;
  }
  const unsigned h = lds_hash(w);
  const unsigned r = h & ((1u << kLdsRBits) - 1);
  unsigned b = h >> kLdsRBits;
  // the home bucket, where almost every lookup ends, then the others in a
  // rolled loop: fifteen unrolled probes per keyed lookup made the
  // expansion loop too large to unroll (a rolled loop from the home bucket
  // on measured 1% slower)
  {
    const int4 q = t[b];
    if (lds_match(q.x, r)) return q.x >> 16;
    if (lds_match(q.y, r)) return q.y >> 16;
    if (lds_match(q.z, r)) return q.z >> 16;
    if (lds_match(q.w, r)) return q.w >> 16;
    if (q4_open(q)) return -1;
    b = (b + 1) & (kLdsBuckets - 1);
  }
#pragma unroll 1
  for (int p = 1; p < kLdsProbe; p++) {
    const int4 q = t[b];
    const unsigned key = r | (unsigned)p << 12;
    if (lds_match(q.x, key)) return q.x >> 16;
    if (lds_match(q.y, key)) return q.y >> 16;
    if (lds_match(q.z, key)) return q.z >> 16;
    if (lds_match(q.w, key)) return q.w >> 16;
    if (q4_open(q)) return -1;
    b = (b + 1) & (kLdsBuckets - 1);
  }
  return -1;
}

// Inserts w (frontier index 0 for now): 1 when this call placed it, with
// its slot in `at` (an int index into t), 0 when it was there, -1 when no
// slot within kLdsProbe buckets was free.  The home bucket's slots are
// claimed in order by compare-and-swap with no read first (at load <= 1/2
// the first swap usually settles it); its swaps keep the fill order, since
// slot k is tried only after slots 0..k-1 were seen taken.  Further buckets
// are read first; a lost race on a slot re-reads the bucket.
template <bool KEYED>
__device__ __forceinline__ int lds_insert(int4 *t, int w, int &at) {
  const unsigned h = lds_hash(w);
  const unsigned r = h & ((1u << kLdsRBits) - 1);
  unsigned b = KEYED ? h >> kLdsRBits : lds_bucket(w);
  // the slot value at probe distance p, and whether slot value x is w's
  auto val = [&](int p) { return KEYED ? (int)(r | (unsigned)p << 12) : w; };
  auto same = [&](int x, int p) { return KEYED ? lds_match(x, r | (unsigned)p << 12) : x == w; };
  {
    int *sl = reinterpret_cast<int *>(&t[b]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = atomicCAS(sl + k, -1, val(0));
      if (x < 0) {
        at = 4 * (int)b + k;
        return 1;
      }
      if (same(x, 0)) return 0;
    }
    b = (b + 1) & (kLdsBuckets - 1);
  }
  for (int p = 1, tries = 0; p < kLdsProbe && tries < 4 * kLdsProbe; tries++) {
    const int4 q = t[b];
    if (same(q.x, p) || same(q.y, p) || same(q.z, p) || same(q.w, p)) return 0;
    int k = q.x < 0 ? 0 : (q.y < 0 ? 1 : (q.z < 0 ? 2 : (q.w < 0 ? 3 : 4)));
    if (k == 4) {
      b = (b + 1) & (kLdsBuckets - 1);
      p++;
      continue;
    }
    const int x = atomicCAS(reinterpret_cast<int *>(&t[b]) + k, -1, val(p));
    if (x < 0) {
      at = 4 * (int)b + k;
      return 1;
    }
    if (same(x, p)) return 0;
  }
  return -1;
}

// Expansion of frontier vertex v at level L (-1 for an idle thread) with
// neighbours q (loaded by the caller).  (Issuing the parents' path-count
// loads ahead of the claims, or expanding two vertices per thread, measured
// slower: 61.4 and 80.4 against 56.4 ms per 2048^2 pass.)
template <bool SIGMA, bool WANT_D, class DT, bool REC, bool STORE = true>
__device__ __forceinline__ void ps_visit_lds(const int4 q, gp<DT> d, gp<u64> sg, gp<ps_rec_t> rec, gp<const u64> lp,
                                             gp<u64> lc, int k, int v, int L, const int4 *tp, const int4 *tl,
                                             int4 *tn, int &ovf, int &bad, u64 &dag, u64 &teps, unsigned *cnt,
                                             PsLds &f, gp<int> spill, gp<int> S, int nbuf, bool bip,
                                             u64 *sv_out = nullptr) {
  const int w[4] = {q.x, q.y, q.z, q.w};
  bool par[4], got[4];
  int pk[4], at[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    par[j] = got[j] = false;
    pk[j] = 0;
    at[j] = 0;
    if (w[j] < 0) continue;
    teps++;
    const int kp = L >= 1 ? lds_find<SIGMA>(tp, w[j]) : -1;
    if (kp >= 0) {
      par[j] = true;
      pk[j] = kp;
    } else if (bip || lds_find<SIGMA>(tl, w[j]) < 0) {
      const int r = lds_insert<SIGMA>(tn, w[j], at[j]);
      got[j] = r > 0;
      bad |= r < 0;
    }
  }
  // the parents' path counts from the previous level's log (every lane
  // loads, entry 0 off the parents)
  u64 x[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (SIGMA) {
      const u64 t = ld_wg(lp + pk[j]);
      x[j] = par[j] ? t : 0ull;
    }
  }
  unsigned claims = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) claims += got[j];
  unsigned pos = wave_reserve(cnt, claims);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!got[j]) continue;
    // the claim's frontier index into its slot (the index bits are 0 until now)
    if (SIGMA) atomicOr(reinterpret_cast<int *>(tn) + at[j], (int)(pos << 16));
    if (S) S[pos] = w[j];
    if (pos < (unsigned)kLdsFront)
      f.fv[nbuf][pos] = w[j];
    else
      spill[pos - kLdsFront] = w[j];
    pos++;
  }
  if (SIGMA) {
    u64 acc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (par[j]) {
        acc += x[j];
        ovf |= acc < x[j];
        dag++;
      }
    }
    if (STORE && L >= 1 && v >= 0) {
      if (REC)
        rec[v] = ps_rec(acc, L);
      else
        sg[v] = acc;
      lc[k] = acc;
    }
    if (!STORE) *sv_out = acc;
  }
  if (STORE && WANT_D && !REC && v >= 0) d[v] = (DT)L;
}

// The level loop over LDS level sets (one workgroup barrier a level).  With
// BC the levels are kept in S (E[l] = start of level l) and a frontier's
// entries beyond the LDS capacity are read back from S; otherwise they go to
// the two `spill` halves.  Returns the deepest level, or -1 when a level set
// overflowed (or, with 16-bit distances, the search went deeper than
// kD16Max levels).  With path counts, lg holds the path-count logs of two
// levels (2 * kLdsLog entries; level L's at lg + (L & 1) * kLdsLog).
// With REC the distances and path counts go to the records rec (ps_rec)
// instead of the rows d and sg.
template <bool SIGMA, bool BC, bool WANT_D, class DT, bool E8, bool REC = false>
__device__ __forceinline__ int ps_levels_lds(const int4 *LE, const u64 *LE8, int n, int s, gp<DT> d, gp<u64> sg,
                                             gp<ps_rec_t> rec, gp<u64> lg, gp<int> spill0, gp<int> S, gp<int> E,
                                             PsLds &f, unsigned *s_cnt, int *s_bad, int &ovf, u64 &teps,
                                             u64 &reached, u64 &dag, bool bip) {
  const int tid = threadIdx.x, T = blockDim.x;
  for (int i = tid; i < 4 * kLdsBuckets; i += T) (&f.tab[0][0])[i] = make_int4(-1, -1, -1, -1);
  if (tid < 3) s_cnt[tid] = 0;
  if (tid == 0) *s_bad = 0;
  __syncthreads();
  if (tid == 0) {
    const unsigned h = lds_hash(s);  // keyed: key r, probe distance 0, frontier index 0
    if (SIGMA)
      f.tab[0][h >> kLdsRBits].x = (int)(h & ((1u << kLdsRBits) - 1));
    else
      f.tab[0][lds_bucket(s)].x = s;
    if (SIGMA) {
      if (REC)
        rec[s] = ps_rec(1, 0);
      else
        sg[s] = 1;
      lg[0] = 1;
    }
    f.fv[0][0] = s;
    if (BC) {
      S[0] = s;
      E[0] = 0;
      E[1] = 1;
    }
  }
  __syncthreads();
  unsigned qn = 1, qb = 0, qs = 1;  // frontier size; its start and end in S
  int L = 0, b = 0, bad = 0;
  reached = 1;
  while (true) {
    unsigned *cnt = &s_cnt[L % 3];
    if (tid == 0) s_cnt[(L + 1) % 3] = 0;
    const int4 *tp = f.tab[(L + 3) & 3], *tl = f.tab[L & 3];
    int4 *tn = f.tab[(L + 1) & 3], *tc = f.tab[(L + 2) & 3];
    gp<const u64> lp = SIGMA ? lg + ((L + 1) & 1) * kLdsLog : nullptr;  // level L-1's log
    gp<u64> lc = SIGMA ? lg + (L & 1) * kLdsLog : nullptr;
    for (int i = tid; i < kLdsBuckets; i += T) tc[i] = make_int4(-1, -1, -1, -1);  // held level L-2
    gp<int> Snext = BC ? S + qs : nullptr;
    gp<const int> sp_cur = BC ? S + qb + kLdsFront : spill0 + (size_t)b * n;
    gp<int> sp_next = BC ? S + qs + kLdsFront : spill0 + (size_t)(b ^ 1) * n;
    // The frontier entries held in LDS first.  No global load feeds a
    // frontier entry in this loop: an entry that might
    // come from HBM made every iteration wait for all of the previous one's
    // global stores (one in-order counter covers loads and stores).
    const unsigned qlds = min(qn, (unsigned)kLdsFront);
    for (unsigned base = 0; base < qlds; base += kLdsSlots * T) {
      // up to kLdsSlots frontier vertices per thread: their table entries
      // are all loaded first (a table entry loaded before the expansion
      // that needs it, and decoded only there, took a 2048^2 pass from 62.5
      // to 56.4 ms), then each vertex is expanded, and their path counts (to
      // the row and to the level's log) and distances are stored after the
      // last expansion.  On this
      // hardware one in-order counter covers loads and stores, so a load
      // waited for after a store also waits for the store: stores between
      // the expansions made each expansion's path-count loads wait for the
      // previous one's stores.
      int vv[kLdsSlots];
      int4 raw[kLdsSlots];
      u64 sv[kLdsSlots];
#pragma unroll
      for (int i = 0; i < kLdsSlots; i++) {
        const unsigned k = base + i * T + tid;
        vv[i] = k < qlds ? f.fv[b][k] : -1;
      }
#pragma unroll
      for (int i = 0; i < kLdsSlots; i++) raw[i] = le_fetch<E8>(LE, LE8, vv[i]);
#pragma unroll
      for (int i = 0; i < kLdsSlots; i++) {
        sv[i] = 0;
        if (base + i * T < qlds)  // uniform
          ps_visit_lds<SIGMA, WANT_D || BC, DT, REC, false>(le_decode<E8>(vv[i], raw[i]), d, sg, rec, lp, lc, 0,
                                                             vv[i], L, tp, tl, tn, ovf, bad, dag, teps, cnt, f,
                                                             sp_next, Snext, b ^ 1, bip, &sv[i]);
      }
#pragma unroll
      for (int i = 0; i < kLdsSlots; i++) {
        if (vv[i] < 0) continue;
        if (REC) {
          // the source's record was written before the first level
          if (L >= 1) {
            rec[vv[i]] = ps_rec(sv[i], L);
            lc[base + i * T + tid] = sv[i];
          }
          continue;
        }
        if (SIGMA && L >= 1) {
          sg[vv[i]] = sv[i];
          lc[base + i * T + tid] = sv[i];
        }
        if (WANT_D || BC) d[vv[i]] = (DT)L;
      }
    }
    // then the entries beyond the LDS frontier (levels larger than it)
    for (unsigned base = qlds; base < qn; base += T) {
      const unsigned k = base + tid;
      const int vs = k < qn ? sp_cur[k - kLdsFront] : -1;
      const int4 q = le_decode<E8>(vs, le_fetch<E8>(LE, LE8, vs));
      ps_visit_lds<SIGMA, WANT_D || BC, DT, REC>(q, d, sg, rec, lp, lc, (int)k, vs, L, tp, tl, tn, ovf, bad, dag,
                                                 teps, cnt, f, sp_next, Snext, b ^ 1, bip);
    }
    if (bad) *s_bad = 1;
    __syncthreads();
    const unsigned q2n = *cnt;
    if (*s_bad) return -1;
    if (q2n == 0) break;
    reached += q2n;
    qn = q2n;
    qb = qs;
    qs += q2n;
    b ^= 1;
    L++;
    if (sizeof(DT) == 2 && L > kD16Max) return -1;
    if (BC && tid == 0) E[L + 1] = (int)qs;
  }
  return L;
}

// ps_writeback for the records of a search with path counts.
__device__ __forceinline__ void ps_writeback_rec(const PsQueue &q, int j, gp<const ps_rec_t> rs, int n) {
  gp<int> dd = glob(q.ddst ? q.ddst[j] : nullptr);
  gp<u64> sd = glob(q.sdst ? q.sdst[j] : nullptr);
  gp<const int> nof = glob(q.nof);
  if (!dd && !sd) return;
  constexpr int U = 6;
  const int T = blockDim.x, tid = threadIdx.x;
  const int steps = (n + U * T - 1) / (U * T);
  auto at = [&](int i, int u) { return i * U * T + u * T + tid; };
  auto cl = [&](int v) { return v < n ? v : n - 1; };
  int xa[U];
  ps_rec_t ra[U];
  {
    int x0[U];
#pragma unroll
    for (int u = 0; u < U; u++) x0[u] = nof[cl(at(0, u))];
#pragma unroll
    for (int u = 0; u < U; u++) xa[u] = nof[cl(at(1, u))];
#pragma unroll
    for (int u = 0; u < U; u++) ra[u] = rs[x0[u]];
  }
  for (int i = 0; i < steps; i++) {
    int xb[U];
    ps_rec_t rb[U];
#pragma unroll
    for (int u = 0; u < U; u++) xb[u] = nof[cl(at(i + 2, u))];
#pragma unroll
    for (int u = 0; u < U; u++) rb[u] = rs[xa[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int v = at(i, u);
      if (v >= n) continue;
      const bool r = ra[u].z >= 0;
      if (dd) dd[v] = r ? ra[u].z : INT_MAX;
      if (sd) sd[v] = r ? ((u64)(unsigned)ra[u].y << 32 | (unsigned)ra[u].x) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      xa[u] = xb[u];
      ra[u] = rb[u];
    }
  }
}

struct PsLdsArgs {
  const int4 *LE;
  int n;
  const int *srcs;    // [nsrc] (cluster-order ids)
  int *const *dist;   // [nsrc] scratch rows (16-bit distances), or null
                      //   (diameter; with a queue: the slot's scratch rows)
  u64 *const *sigma;  // [nsrc] with path counts: record rows (ps_rec, 16
                      //   bytes a vertex; with a queue: the slot's scratch
                      //   records at q.sscr + slot * 2n), or null
  int *queues;        // [grid][2n+2] frontier spill
  int *depth;
  u64 *part;          // [nsrc][4], as PsArgs; part[3] bit 1: level set overflow
  int *ovf;
  int flags;          // kLds* graph facts
  const u64 *LE8;     // or null: the neighbour table as int16 deltas
  PsQueue q;          // source queue and in-kernel write-back
  u64 *slog;          // with path counts: [grid][2 * kLdsLog] path-count logs
};

// The search: with distance rows only (WANT_D) the row, in 16 bits
// (kD16None: not reached), is first set to kD16None, the only row
// initialised: it tells reached vertices apart when the rows are written
// back (k_ps_unpermute_ds).  With path counts the records (ps_rec) are set
// to distance -1 first and written for reached vertices
// (k_ps_unpermute_rec).  On a connected graph every vertex is written by
// the search, so the fill is skipped.  A source deeper than kD16Max levels
// stops as a level-set overflow does (the host reruns the launch).
template <bool SIGMA, bool WANT_D, bool E8>
__global__ __launch_bounds__(1024) void k_ps_lds(PsLdsArgs a) {
  constexpr bool REC = SIGMA && WANT_D;
  __shared__ PsLds f;
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf, s_bad, s_job;
  const int tid = threadIdx.x, T = blockDim.x;
  const size_t slot = blockIdx.x, n = (size_t)a.n;
  gp<int> spill = glob(a.queues + slot * (2 * n + 2));
  for (int j = ps_next_job(a.q, &s_job, true); j < a.q.nsrc; j = ps_next_job(a.q, &s_job, false)) {
    const int s = a.srcs[j];
    gp<unsigned short> d = nullptr;
    if (WANT_D && !REC)
      d = glob(reinterpret_cast<unsigned short *>(a.dist && a.dist[j] ? a.dist[j] : a.q.dscr + slot * n));
    gp<ps_rec_t> rec = nullptr;
    if (REC) rec = glob(reinterpret_cast<ps_rec_t *>(a.sigma[j] ? a.sigma[j] : a.q.sscr + slot * 2 * n));
    gp<u64> lg = glob(SIGMA ? a.slog + slot * (2 * kLdsLog) : nullptr);
    if (!(a.flags & kLdsConnected)) {
      if (REC)
        for (int i = tid; i < a.n; i += T) rec[i] = ps_rec(0, -1);
      else if (WANT_D)
        for (int i = tid; i < a.n; i += T) d[i] = kD16None;
    }
    if (tid == 0) s_ovf = 0;
    __syncthreads();
    int ovf = 0;
    u64 teps = 0, reached = 1, dag = 0;
    const int L = ps_levels_lds<SIGMA, false, WANT_D, unsigned short, E8, REC>(
        a.LE, a.LE8, a.n, s, d, nullptr, rec, lg, spill, nullptr, nullptr, f, s_cnt, &s_bad, ovf, teps, reached, dag,
        (a.flags & kLdsBipartite) != 0);
    ps_finish(teps, dag, reached, L < 0 ? 0 : L, ovf, &s_ovf, s_red, a.part, a.depth, a.ovf, j);
    if (tid == 0 && L < 0) a.part[4 * (size_t)j + 3] |= 2ull;
    // an overflowed source is searched again by the host, rows included
    if (REC && a.q.nof && L >= 0) ps_writeback_rec(a.q, j, rec, a.n);
    if (WANT_D && !REC && a.q.nof && L >= 0) ps_writeback<unsigned short>(a.q, j, d, nullptr, a.n, kD16None);
    __syncthreads();  // rows and s_job read before the next source
  }
}

// Brandes over LDS level sets: the forward above with the BFS order in S
// and every reached vertex's distance written, then the backward of k_ps_bc
// (successors d = l+1; on an undirected graph every neighbour of a reached
// vertex is reached, so only written distances are read).  The dependency
// row is set to 0 first (the BC reduce and write-back read every vertex).
template <bool E8>
__global__ __launch_bounds__(1024) void k_ps_bc_lds(PsBcArgs a) {
  __shared__ PsLds f;
  __shared__ unsigned s_cnt[3];
  __shared__ u64 s_red[16][2];
  __shared__ int s_ovf, s_bad;
  const int tid = threadIdx.x, T = blockDim.x;
  const int s = a.srcs[blockIdx.x];
  gp<int> d = glob(a.dist[blockIdx.x]);
  gp<u64> sg = glob(a.sigma[blockIdx.x]);
  gp<float> dl = glob(a.delta[blockIdx.x]);
  gp<float> ql = glob(a.quot[blockIdx.x]);
  gp<int> S = glob(a.stack + (size_t)blockIdx.x * (2 * (size_t)a.n + 2));
  gp<int> E = S + a.n;
  for (int i = tid; i < a.n; i += T) dl[i] = 0.0f;
  if (tid == 0) s_ovf = 0;
  __syncthreads();
  int ovf = 0;
  u64 teps = 0, reached = 1, dag = 0;
  gp<u64> lg = glob(a.slog + (size_t)blockIdx.x * (2 * kLdsLog));
  const int L = ps_levels_lds<true, true, true, int, E8>(a.LE, a.LE8, a.n, s, d, sg, nullptr, lg, nullptr, S, E, f,
                                                         s_cnt, &s_bad, ovf, teps, reached, dag,
                                                         (a.lds_flags & kLdsBipartite) != 0);
  __syncthreads();
  if (L < 0) {  // rerun by the host on k_ps_bc
    ps_finish(teps, dag, reached, 0, ovf, &s_ovf, s_red, a.part, a.depth, a.ovf);
    if (tid == 0) a.part[4 * blockIdx.x + 3] |= 2ull;
    return;
  }
  for (int i = E[L] + tid; L >= 1 && i < E[L + 1]; i += T) {
    const int v = S[i];
    ql[v] = 1.0f / (float)ld_wg(sg + v);
  }
  __syncthreads();
  for (int l = L - 1; l >= 1; l--) {
    for (int i = E[l] + tid; i < E[l + 1]; i += T) {
      const int v = S[i];
      const float sv = (float)ld_wg(sg + v);
      const int4 q4 = le_decode<E8>(v, le_fetch<E8>(a.LE, a.LE8, v));
      const int w[4] = {q4.x, q4.y, q4.z, q4.w};
      bool on[4];
      float qw[4];
      ps_successors<4>(d, ql, w, l, on, qw);
      float sum = 0.0f, comp = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (on[j]) ps_neumaier(sum, comp, sv * qw[j]);
      const float dv = sum + comp;
      dl[v] = dv;
      ql[v] = (1.0f + dv) / sv;
    }
    __syncthreads();
  }
  ps_finish(teps, dag, reached, L, ovf, &s_ovf, s_red, a.part, a.depth, a.ovf);
}

// Both rows of an LDS level-set search back to the caller's order in one
// pass: dd[b][v] = dist (INT_MAX where the 16-bit row says not reached),
// sd[b][v] = path count where reached, else 0; the distance destination
// may be null, and so may the path counts (ssrc and sdst both null: a
// search for distances only).
// blockIdx.y selects the row.  XCD-aware: blocks are dispatched to the 8
// XCDs round-robin, so block b takes contiguous chunk (b % 8) * (G / 8) +
// b / 8 of the row's G chunks (G = gridDim.x, a multiple of 8) and each
// XCD gathers from one eighth of the vertex range; the clusters a chunk
// gathers from are shared with the neighbouring chunks of the same XCD
// (and L2) instead of being fetched by every XCD.
__global__ __launch_bounds__(256) void k_ps_unpermute_ds(int *const *dsrc, u64 *const *ssrc, int *const *ddst,
                                                         u64 *const *sdst, const int *nof, int n) {
  gp<int> dd = glob(ddst ? ddst[blockIdx.y] : nullptr);
  gp<u64> sd = glob(sdst ? sdst[blockIdx.y] : nullptr);
  gp<const unsigned short> ds = glob(reinterpret_cast<const unsigned short *>(dsrc[blockIdx.y]));
  gp<const u64> ss = glob(ssrc ? (const u64 *)ssrc[blockIdx.y] : nullptr);
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned c = (b % 8) * (G / 8) + b / 8;
  const size_t chunk = ((size_t)n + G - 1) / G;
  const size_t lo = (size_t)c * chunk, hi = min((size_t)n, lo + chunk);
  gp<const int> nf = glob(nof);
  // sixteen vertices per thread with every load issued before the first
  // store (the map entries, then the gathers): one wait per sixteen.  On
  // the 2048^2 grid's 256-source pass, 16 took the pass from 40.4 to 38.9
  // ms against 4 (2: 40.7, 8: 40.5, 24: 41.2, 32: 39.1; two steps of 16
  // per thread: 39.9)
  constexpr int U = 16;
  const size_t T = blockDim.x;
  size_t v0 = lo + threadIdx.x;
  for (; v0 + (U - 1) * T < hi; v0 += U * T) {
    int x[U];
    unsigned short dv[U];
    u64 sv[U];
#pragma unroll
    for (int u = 0; u < U; u++) x[u] = nf[v0 + u * T];
#pragma unroll
    for (int u = 0; u < U; u++) {
      dv[u] = ds[x[u]];
      sv[u] = ss ? ss[x[u]] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (dd) dd[v0 + u * T] = dv[u] != kD16None ? (int)dv[u] : INT_MAX;
      if (sd) sd[v0 + u * T] = dv[u] != kD16None ? sv[u] : 0ull;
    }
  }
  for (size_t v = v0; v < hi; v += T) {
    const int x = nf[v];
    const unsigned short dv = ds[x];
    const u64 sv = ss ? ss[x] : 0ull;
    if (dd) dd[v] = dv != kD16None ? (int)dv : INT_MAX;
    if (sd) sd[v] = dv != kD16None ? sv : 0ull;
  }
}

// The same for the records of a search with path counts (ps_rec; distance
// -1 where not reached).
__global__ __launch_bounds__(256) void k_ps_unpermute_rec(const ps_rec_t *const *rsrc, int *const *ddst,
                                                          u64 *const *sdst, const int *nof, int n) {
  gp<int> dd = glob(ddst ? ddst[blockIdx.y] : nullptr);
  gp<u64> sd = glob(sdst ? sdst[blockIdx.y] : nullptr);
  gp<const ps_rec_t> rs = glob(rsrc[blockIdx.y]);
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned c = (b % 8) * (G / 8) + b / 8;
  const size_t chunk = ((size_t)n + G - 1) / G;
  const size_t lo = (size_t)c * chunk, hi = min((size_t)n, lo + chunk);
  gp<const int> nf = glob(nof);
  constexpr int U = 16;
  const size_t T = blockDim.x;
  size_t v0 = lo + threadIdx.x;
  for (; v0 + (U - 1) * T < hi; v0 += U * T) {
    int x[U];
    ps_rec_t r[U];
#pragma unroll
    for (int u = 0; u < U; u++) x[u] = nf[v0 + u * T];
#pragma unroll
    for (int u = 0; u < U; u++) r[u] = rs[x[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool ok = r[u].z >= 0;
      if (dd) dd[v0 + u * T] = ok ? r[u].z : INT_MAX;
      if (sd) sd[v0 + u * T] = ok ? ((u64)(unsigned)r[u].y << 32 | (unsigned)r[u].x) : 0ull;
    }
  }
  for (size_t v = v0; v < hi; v += T) {
    const ps_rec_t r = rs[nf[v]];
    const bool ok = r.z >= 0;
    if (dd) dd[v] = ok ? r.z : INT_MAX;
    if (sd) sd[v] = ok ? ((u64)(unsigned)r.y << 32 | (unsigned)r.x) : 0ull;
  }
}

}  // namespace gu
