// gu_kernels.hpp -- device side of the batched multi-source search engine.
//
// One launch sequence serves a batch of up to 64*NW sources ("lanes").  Each
// vertex carries NW 64-bit masks: visit (lanes that reached it), and three
// rotating frontier slots (lanes for which it sits on the current frontier,
// the next frontier, and the stale one being cleared).  A level is
//   expand    top-down (one of the six reference expansion strategies) or
//             bottom-up (pull from the frontier, with the path-count pull
//             fused in)
//   finalize  top-down only: commit the newly reached lanes of every vertex
//             pushed this level and pull its path counts (lane = source)
//   control   one wave: termination, queue rotation, direction choice
// so distances and path counts are computed for all lanes by one pass over
// each edge per level instead of one pass per source (the reference runs one
// thread block per source: shuffle_based.cuh:104).
//
// Per-vertex per-lane values (distance, path count, dependency) are stored
// vertex-major [n][64*NW] so that a wave reading one vertex's row for all of
// its lanes issues one contiguous 256/512-byte access; output kernels
// transpose them to the reference's source-major rows.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

#include "gu_wave.hpp"

namespace gu {

using u64 = unsigned long long;

enum Variant { V_LINEAR = 0, V_EDGE_PARALLEL = 1, V_WARP = 2, V_SCAN = 3, V_CTA_WARP = 4, V_SHUFFLE = 5 };

struct Ctrl {
  int level;           // frontier depth L of the level being run
  int done;
  int mode;            // 0 top-down, 1 bottom-up (for `level`)
  int last_level;      // deepest non-empty level once done
  unsigned qlen[3];    // queue length per slot (level % 3)
  unsigned work[4];    // dynamic work counters, zeroed every level
  u64 fr_edges[3];     // sum of out-degrees of each slot's queue
  u64 unvis_edges;     // sum of out-degrees of vertices some lane has not reached
  u64 unvis_dec;       // decrements gathered during the level
  u64 teps;            // sum over lanes of out-degrees of reached vertices
  u64 reached;         // sum over lanes of reached vertices
  u64 dag;             // sum over lanes of shortest-path DAG edges pulled
  long long S_len;     // stack mode: end of the current queue in S
  int overflow;        // a path count wrapped mod 2^64
  int td_levels, bu_levels;
  int direction, alpha, beta;
  int n;
  int pad[3];
};

template <int NW>
struct Eng {
  const int *R;
  const int *C;
  int n;
  int nb;              // active lanes
  u64 *visit;          // [n][NW]
  u64 *fr;             // [3][n][NW]
  int *q;              // BFS: [3][n] queue slots; stack mode: S
  u64 *smask;          // stack mode: lanes of each S entry [S_cap][NW]
  long long *ep;       // stack mode: level start offsets in S
  int *tag;            // [n] claim tags (NW > 1 pushes, source dedup)
  int *dvm;            // [n][64*NW] distances
  u64 *svm;            // [n][64*NW] path counts
  Ctrl *ctrl;
  int *off;            // edge-parallel: frontier degree offsets [n+1]
  long long *tsum;     // scan tile sums
  const int *srcs;     // [nb] batch sources
  const int2 *blocks;  // bottom-up vertex blocks: [v0, v1), <= 64 vertices,
  int nblocks;         //   bounded edge count, hubs excluded
  const int4 *segs;    // hub adjacency segments: (w, e0, e1, hub index)
  int nsegs;
  const int *hubs;     // vertices of degree > hub_deg
  int nhubs;
  int hub_deg;
  u64 *hub_found;      // [nhubs][NW] lanes found by the hub's segments
  u64 all[NW];         // mask of active lanes per word
  int stack;           // keep every level's queue (Brandes backward needs it)
  int sigma;           // compute path counts
};

constexpr int kWL = 64;  // lanes per mask word

__device__ __forceinline__ int slot_of(int L) { return L % 3; }

template <int NW>
__device__ __forceinline__ int *qptr(const Eng<NW> &e, int L) {
  return e.stack ? e.q + e.ep[L] : e.q + (size_t)slot_of(L) * e.n;
}
template <int NW>
__device__ __forceinline__ u64 *frs(const Eng<NW> &e, int L) {
  return e.fr + (size_t)slot_of(L) * e.n * NW;
}

__device__ __forceinline__ unsigned bcast(unsigned x) {
  return (unsigned)__builtin_amdgcn_readfirstlane((int)x);
}

// Top-down visit of edge v->w for the lanes f of v (shuffle_based.cu:60-77,
// count_shortest_paths' atomicCAS visit): lanes of f that have not reached w
// are OR-ed into w's next-frontier mask.  Returns true for exactly one
// caller per vertex per level: the one that must push w.
template <int NW>
__device__ __forceinline__ bool discover(const Eng<NW> &e, int L, const u64 *f, int w) {
  u64 *fn = frs(e, L + 1) + (size_t)w * NW;
  const u64 *vis = e.visit + (size_t)w * NW;
  if (NW == 1) {
    u64 m = f[0] & ~vis[0];
    if (!m) return false;
    u64 cur = fn[0];
    if ((cur & m) == m) return false;  // already set, so already pushed
    u64 old = atomicOr(fn, m);
    return old == 0;
  } else {
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      u64 m = f[j] & ~vis[j];
      if (m) {
        any = true;
        if ((fn[j] & m) != m) atomicOr(fn + j, m);
      }
    }
    if (!any) return false;
    if (e.tag[w] >= L + 1) return false;
    return atomicMax(&e.tag[w], L + 1) < L + 1;
  }
}

// Wave-aggregated push of w into the next level's queue: one atomic per wave.
template <int NW>
__device__ __forceinline__ void push_wave(const Eng<NW> &e, int L, bool p, int w) {
  uint64_t m = ballot(p);
  if (!m) return;
  int leader = __ffsll((unsigned long long)m) - 1;
  unsigned base = 0;
  if (lane_id() == leader) base = atomicAdd(&e.ctrl->qlen[slot_of(L + 1)], (unsigned)__popcll(m));
  base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
  if (p) qptr(e, L + 1)[base + rank_in(m)] = w;
}

template <int NW>
__device__ __forceinline__ void load_mask(const u64 *p, u64 *f) {
#pragma unroll
  for (int j = 0; j < NW; j++) f[j] = p[j];
}

// Strip-mine the adjacency list [r0,r1) of a frontier vertex with masks f,
// 64 entries per step (uniform trip count across the wave).
template <int NW>
__device__ __forceinline__ void strip_wave(const Eng<NW> &e, int L, const u64 *f, int r0, int r1) {
  const int lane = lane_id();
  for (int e0 = r0; e0 < r1; e0 += kWave) {
    int ed = e0 + lane;
    bool valid = ed < r1;
    int w = valid ? e.C[ed] : 0;
    bool p = valid && discover<NW>(e, L, f, w);
    push_wave<NW>(e, L, p, w);
  }
}

// ---- top-down strategies ------------------------------------------------

// shuffle_based.cuh:192-223: one wave per frontier vertex, waves fetch the
// next queue entries dynamically; lanes stride the adjacency list.
template <int NW>
__device__ void td_shuffle(const Eng<NW> &e, int L) {
  const int lane = lane_id();
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const u64 *fc = frs(e, L);
  constexpr unsigned CH = 4;
  while (true) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&e.ctrl->work[0], CH);
    base = bcast(base);
    if (base >= qn) break;
    unsigned end = base + CH < qn ? base + CH : qn;
    for (unsigned k = base; k < end; k++) {
      int v = Q[k];
      u64 f[NW];
      load_mask<NW>(fc + (size_t)v * NW, f);
      strip_wave<NW>(e, L, f, e.R[v], e.R[v + 1]);
    }
  }
}

// warp_based.cu:89-166: every lane holds one frontier vertex; the wave
// repeatedly elects the highest lane that still has adjacency left
// (race_and_resolve) and strip-mines that lane's list.
template <int NW>
__device__ void td_warp_core(const Eng<NW> &e, int L, int v, int r, int re) {
  const int lane = lane_id();
  const u64 *fc = frs(e, L);
  while (true) {
    bool pend = r < re;
    if (!ballot(pend)) break;
    int win = race_and_resolve_wave(pend);
    int rw = readlane(r, win), rew = readlane(re, win), vw = readlane(v, win);
    if (lane == win) r = re;  // the same lane cannot win twice
    u64 f[NW];
    load_mask<NW>(fc + (size_t)vw * NW, f);
    strip_wave<NW>(e, L, f, rw, rew);
  }
}

template <int NW>
__device__ void td_warp(const Eng<NW> &e, int L) {
  const int lane = lane_id();
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const unsigned waves = gridDim.x * (blockDim.x / kWave);
  const unsigned gw = blockIdx.x * (blockDim.x / kWave) + wave_id();
  for (unsigned base = gw * kWave; base < qn; base += waves * kWave) {
    unsigned k = base + lane;
    bool have = k < qn;
    int v = have ? Q[k] : 0;
    int r = have ? e.R[v] : 0;
    int re = have ? e.R[v + 1] : 0;
    td_warp_core<NW>(e, L, v, r, re);
  }
}

// CTA_warp.cu:105-186 then :188-265: vertices whose degree exceeds the
// threshold are expanded by the whole workgroup, one at a time; the rest by
// the wave election above.
template <int NW>
__device__ void td_cta_warp(const Eng<NW> &e, int L) {
  constexpr int kThreshold = 2048;  // CTA_warp.cu:44
  __shared__ int s_win, s_v, s_r, s_re;
  const int tid = threadIdx.x;
  const int T = blockDim.x;
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const u64 *fc = frs(e, L);
  for (unsigned tile = blockIdx.x; (size_t)tile * T < qn; tile += gridDim.x) {
    unsigned k = tile * T + tid;
    bool have = k < qn;
    int v = have ? Q[k] : 0;
    int r = have ? e.R[v] : 0;
    int re = have ? e.R[v + 1] : 0;
    while (true) {
      if (tid == 0) s_win = -1;
      __syncthreads();
      if (re - r > kThreshold) atomicMax(&s_win, tid);
      __syncthreads();
      int win = s_win;
      if (win < 0) break;
      if (tid == win) {
        s_v = v;
        s_r = r;
        s_re = re;
        r = re;
      }
      __syncthreads();
      int vw = s_v, rw = s_r, rew = s_re;
      u64 f[NW];
      load_mask<NW>(fc + (size_t)vw * NW, f);
      for (int e0 = rw; e0 < rew; e0 += T) {
        int ed = e0 + tid;
        bool valid = ed < rew;
        int w = valid ? e.C[ed] : 0;
        bool p = valid && discover<NW>(e, L, f, w);
        push_wave<NW>(e, L, p, w);
      }
      __syncthreads();
    }
    td_warp_core<NW>(e, L, v, r, re);
  }
}

// scan_based.cu:105-169: a workgroup scans the degrees of its frontier
// tile, then gathers the tile's edges blockDim at a time through LDS.
template <int NW>
__device__ void td_scan(const Eng<NW> &e, int L) {
  __shared__ int s_e[1024];
  __shared__ int s_slot[1024];
  __shared__ int s_vert[1024];
  __shared__ int s_scan[17];
  const int tid = threadIdx.x;
  const int T = blockDim.x;
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const u64 *fc = frs(e, L);
  for (unsigned tile = blockIdx.x; (size_t)tile * T < qn; tile += gridDim.x) {
    unsigned k = tile * T + tid;
    bool have = k < qn;
    int v = have ? Q[k] : 0;
    int r = have ? e.R[v] : 0;
    int re = have ? e.R[v + 1] : 0;
    s_vert[tid] = v;
    int total;
    int rsv = block_exclusive_sum(re - r, s_scan, &total);
    for (int progress = 0; progress < total; progress += T) {
      while (rsv < progress + T && r < re) {
        s_e[rsv - progress] = r;
        s_slot[rsv - progress] = tid;
        rsv++;
        r++;
      }
      __syncthreads();
      int cnt = total - progress < T ? total - progress : T;
      bool p = false;
      int w = 0;
      if (tid < cnt) {
        int ed = s_e[tid];
        int vv = s_vert[s_slot[tid]];
        u64 f[NW];
        load_mask<NW>(fc + (size_t)vv * NW, f);
        w = e.C[ed];
        p = discover<NW>(e, L, f, w);
      }
      push_wave<NW>(e, L, p, w);
      __syncthreads();
    }
  }
}

// linear_with_atomics.cu:120-157: one thread per frontier vertex walking its
// adjacency list serially; pushes are single atomics.
template <int NW>
__device__ void td_linear(const Eng<NW> &e, int L) {
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const u64 *fc = frs(e, L);
  int *Qn = qptr(e, L + 1);
  const unsigned nt = gridDim.x * blockDim.x;
  for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < qn; k += nt) {
    int v = Q[k];
    u64 f[NW];
    load_mask<NW>(fc + (size_t)v * NW, f);
    for (int ed = e.R[v]; ed < e.R[v + 1]; ed++) {
      int w = e.C[ed];
      if (discover<NW>(e, L, f, w)) {
        unsigned pos = atomicAdd(&e.ctrl->qlen[slot_of(L + 1)], 1u);
        Qn[pos] = w;
      }
    }
  }
}

// Index of the last segment whose offset is <= x within off[lo..hi]
// (upper_bound - 1), off non-decreasing and off[lo] <= x.
__device__ __forceinline__ int seg_search(const int *off, int lo, int hi, long long x) {
  int a = lo + 1, b = hi + 1;  // first index in [lo+1, hi+1) with off > x
  while (a < b) {
    int mid = a + ((b - a) >> 1);
    if ((long long)off[mid] <= x)
      a = mid + 1;
    else
      b = mid;
  }
  return a - 1;
}

// edge_parallel.cu:236-247 restated over the frontier's edges only: the
// frontier's degree offsets (device-wide scan) are merged with the edge
// indices, each workgroup taking an equal slice of edges (merge-path load
// balancing; the reference scans all m edges every level instead).
template <int NW>
__device__ void td_edge_parallel(const Eng<NW> &e, int L) {
  constexpr int IPT = 4;
  constexpr int kCap = 1024 * IPT + 1;
  __shared__ int s_off[kCap];
  __shared__ int s_lo, s_hi;
  const int tid = threadIdx.x;
  const int T = blockDim.x;
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  if (qn == 0) return;
  const u64 *fc = frs(e, L);
  const long long E = e.off[qn];
  const long long TE = (long long)T * IPT;
  for (long long tile = blockIdx.x; tile * TE < E; tile += gridDim.x) {
    long long e0 = tile * TE;
    long long e1 = e0 + TE < E ? e0 + TE : E;
    if (tid == 0) {
      s_lo = seg_search(e.off, 0, (int)qn - 1, e0);
      s_hi = seg_search(e.off, s_lo, (int)qn - 1, e1 - 1);
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    const bool in_lds = hi - lo + 1 <= kCap;
    if (in_lds)
      for (int i = tid; i <= hi - lo; i += T) s_off[i] = e.off[lo + i];
    __syncthreads();
    for (int it = 0; it < IPT; it++) {
      long long ed = e0 + (long long)it * T + tid;
      bool valid = ed < e1;
      int w = 0;
      bool p = false;
      if (valid) {
        int vi;
        if (in_lds)
          vi = lo + seg_search(s_off, 0, hi - lo, ed);
        else
          vi = seg_search(e.off, lo, hi, ed);
        int v = Q[vi];
        u64 f[NW];
        load_mask<NW>(fc + (size_t)v * NW, f);
        w = e.C[e.R[v] + (int)(ed - e.off[vi])];
        p = discover<NW>(e, L, f, w);
      }
      push_wave<NW>(e, L, p, w);
    }
    __syncthreads();
  }
}

// Add the path-count rows of the predecessor lanes selected by pm (lane k
// holds neighbour v with lane mask mv): acc[b] += sigma[v_k][b] for every b in
// mv_k.  Four rows are loaded before any is consumed so each wave keeps four
// row loads in flight.
template <int NW>
__device__ __forceinline__ void add_rows(const Eng<NW> &e, int v, const u64 *mv, uint64_t pm, u64 *acc,
                                         int *ovf, u64 *dag) {
  const int lane = lane_id();
  while (pm) {
    int k[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (pm) {
        k[u] = __ffsll((unsigned long long)pm) - 1;
        pm &= pm - 1;
      } else {
        k[u] = -1;
      }
    }
    u64 x[4][NW];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (k[u] >= 0) {
        const int vk = readlane(v, k[u]);
        const u64 *row = e.svm + (size_t)vk * (kWL * NW);
#pragma unroll
        for (int j = 0; j < NW; j++) {
          const u64 mk = readlane64(mv[j], k[u]);
          *dag += (u64)__popcll(mk);
          x[u][j] = ((mk >> lane) & 1ull) ? row[j * kWL + lane] : 0ull;
        }
      } else {
#pragma unroll
        for (int j = 0; j < NW; j++) x[u][j] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
#pragma unroll
      for (int j = 0; j < NW; j++) {
        u64 sum = acc[j] + x[u][j];
        *ovf |= sum < x[u][j];
        acc[j] = sum;
      }
    }
  }
}

// ---- path-count pull (lane = source) -------------------------------------

// For vertex w and target lanes tgt (lanes that reach w at level L+1),
// accumulate over w's neighbours v: for each lane b set in fr_L[v] & tgt,
// acc[b] += sigma[v][b].  Returns (in found) the OR of fr_L[v] & tgt over
// all neighbours.  Without path counts it stops as soon as every target lane
// has been found (bottom-up early exit).
template <int NW, bool SIGMA>
__device__ __forceinline__ void pull(const Eng<NW> &e, const u64 *fc, int r0, int r1,
                                     const u64 *tgt, u64 *found, u64 *acc, int *ovf, u64 *dag) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < NW; j++) {
    found[j] = 0;
    acc[j] = 0;
  }
  for (int e0 = r0; e0 < r1; e0 += kWave) {
    int ed = e0 + lane;
    bool valid = ed < r1;
    int v = valid ? e.C[ed] : 0;
    u64 mv[NW];
    bool has = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      mv[j] = valid ? (fc[(size_t)v * NW + j] & tgt[j]) : 0;
      has |= mv[j] != 0;
    }
    if (SIGMA) {
#pragma unroll
      for (int j = 0; j < NW; j++) found[j] |= mv[j];
      add_rows<NW>(e, v, mv, ballot(has), acc, ovf, dag);
    } else {
      bool full = true;
#pragma unroll
      for (int j = 0; j < NW; j++) {
        found[j] |= mv[j];
        u64 all_found = wave_or(found[j]);
        if ((all_found & tgt[j]) != tgt[j]) full = false;
      }
      if (full) break;
    }
  }
#pragma unroll
  for (int j = 0; j < NW; j++) found[j] = wave_or(found[j]);
}

// Commit lanes nw of vertex w reached at level L+1 (lane b writes its own
// distance / path count).
template <int NW, bool SIGMA>
__device__ __forceinline__ void commit(const Eng<NW> &e, int L, int w, const u64 *vo,
                                       const u64 *nw, const u64 *acc) {
  const int lane = lane_id();
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NW; j++) e.visit[(size_t)w * NW + j] = vo[j] | nw[j];
  }
  int *drow = e.dvm + (size_t)w * (kWL * NW);
  u64 *srow = e.svm + (size_t)w * (kWL * NW);
#pragma unroll
  for (int j = 0; j < NW; j++) {
    if ((nw[j] >> lane) & 1ull) {
      drow[j * kWL + lane] = L + 1;
      if (SIGMA) srow[j * kWL + lane] = acc[j];
    }
  }
}

template <int NW>
__device__ __forceinline__ bool full_mask(const Eng<NW> &e, const u64 *vo, const u64 *nw) {
  bool f = true;
#pragma unroll
  for (int j = 0; j < NW; j++) f &= (vo[j] | nw[j]) == e.all[j];
  return f;
}

// Bottom-up level: every vertex that some lane has not reached pulls from
// its neighbours on the frontier.  Waves take precomputed blocks of at most
// 64 consecutive vertices with a bounded edge count (hubs are left to the
// segment pass); a block's discoveries are pushed with one atomic.
template <int NW, bool SIGMA>
__device__ void bu_level(const Eng<NW> &e, int L) {
  const int lane = lane_id();
  const u64 *fc = frs(e, L);
  u64 *fn = frs(e, L + 1);
  int *Qn = qptr(e, L + 1);
  u64 fe_acc = 0, ud_acc = 0, dag = 0;
  int ovf = 0;
  while (true) {
    unsigned bi = 0;
    if (lane == 0) bi = atomicAdd(&e.ctrl->work[1], 1u);
    bi = bcast(bi);
    if (bi >= (unsigned)e.nblocks) break;
    const int2 blk = e.blocks[bi];
    const int nblk = blk.y - blk.x;
    uint64_t disc = 0;
    u64 mine[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) mine[j] = 0;
    for (int i = 0; i < nblk; i++) {
      const int w = blk.x + i;
      const int r0 = e.R[w], r1 = e.R[w + 1];
      if (r0 == r1) continue;
      u64 vo[NW], tgt[NW];
      bool any = false;
#pragma unroll
      for (int j = 0; j < NW; j++) {
        vo[j] = e.visit[(size_t)w * NW + j];
        tgt[j] = e.all[j] & ~vo[j];
        any |= tgt[j] != 0;
      }
      if (!any) continue;
      u64 found[NW], acc[NW];
      pull<NW, SIGMA>(e, fc, r0, r1, tgt, found, acc, &ovf, &dag);
      bool anyn = false;
#pragma unroll
      for (int j = 0; j < NW; j++) anyn |= found[j] != 0;
      if (!anyn) continue;
      commit<NW, SIGMA>(e, L, w, vo, found, acc);
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++) fn[(size_t)w * NW + j] = found[j];
      }
      if (lane == i) {
#pragma unroll
        for (int j = 0; j < NW; j++) mine[j] = found[j];
      }
      disc |= 1ull << i;
      fe_acc += (u64)(r1 - r0);
      if (full_mask<NW>(e, vo, found)) ud_acc += (u64)(r1 - r0);
    }
    if (disc) {
      unsigned base = 0;
      if (lane == 0) base = atomicAdd(&e.ctrl->qlen[slot_of(L + 1)], (unsigned)__popcll(disc));
      base = bcast(base);
      if ((disc >> lane) & 1ull) {
        unsigned pos = base + (unsigned)rank_in(disc);
        Qn[pos] = blk.x + lane;
        if (e.stack) {
#pragma unroll
          for (int j = 0; j < NW; j++) e.smask[((size_t)e.ep[L + 1] + pos) * NW + j] = mine[j];
        }
      }
    }
  }
  if (ballot(ovf != 0) && lane == 0) atomicOr(&e.ctrl->overflow, 1);
  if (lane == 0) {
    if (fe_acc) atomicAdd(&e.ctrl->fr_edges[slot_of(L + 1)], fe_acc);
    if (ud_acc) atomicAdd(&e.ctrl->unvis_dec, ud_acc);
    if (dag) atomicAdd(&e.ctrl->dag, dag);
  }
}

// Top-down finalize: for every non-hub vertex pushed this level, commit its
// new lanes and (with path counts) pull sigma from its frontier neighbours.
// Hubs only record their stack lanes here; their pull runs in segments.
template <int NW, bool SIGMA>
__device__ void td_finalize(const Eng<NW> &e, int L) {
  constexpr int CH = 16;
  const int lane = lane_id();
  const int *Qn = qptr(e, L + 1);
  const unsigned qn = e.ctrl->qlen[slot_of(L + 1)];
  const u64 *fc = frs(e, L);
  const u64 *fn = frs(e, L + 1);
  u64 fe_acc = 0, ud_acc = 0, dag = 0;
  int ovf = 0;
  while (true) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&e.ctrl->work[2], (unsigned)CH);
    base = bcast(base);
    if (base >= qn) break;
    const int cnt = qn - base < (unsigned)CH ? (int)(qn - base) : CH;
    const int wl = lane < cnt ? Qn[base + lane] : 0;
    for (int i = 0; i < cnt; i++) {
      const int w = readlane(wl, i);
      u64 vo[NW], nw[NW], acc[NW];
#pragma unroll
      for (int j = 0; j < NW; j++) {
        vo[j] = e.visit[(size_t)w * NW + j];
        nw[j] = fn[(size_t)w * NW + j];
        acc[j] = 0;
      }
      if (e.stack && lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++) e.smask[((size_t)e.ep[L + 1] + base + i) * NW + j] = nw[j];
      }
      const int r0 = e.R[w], r1 = e.R[w + 1];
      if (r1 - r0 > e.hub_deg) continue;  // committed by the hub pass
      if (SIGMA) {
        u64 found[NW];
        pull<NW, true>(e, fc, r0, r1, nw, found, acc, &ovf, &dag);
      }
      commit<NW, SIGMA>(e, L, w, vo, nw, acc);
      fe_acc += (u64)(r1 - r0);
      if (full_mask<NW>(e, vo, nw)) ud_acc += (u64)(r1 - r0);
    }
  }
  if (ballot(ovf != 0) && lane == 0) atomicOr(&e.ctrl->overflow, 1);
  if (lane == 0) {
    if (fe_acc) atomicAdd(&e.ctrl->fr_edges[slot_of(L + 1)], fe_acc);
    if (ud_acc) atomicAdd(&e.ctrl->unvis_dec, ud_acc);
    if (dag) atomicAdd(&e.ctrl->dag, dag);
  }
}

// Hub segments (both directions): each wave pulls one segment of a hub's
// adjacency for the hub's target lanes (top-down: the lanes pushed this
// level; bottom-up: the lanes not yet reached) and adds its partial path
// counts with 64-bit atomics -- exact, so the sum does not depend on the
// order of the segments.  Bottom-up also ORs the lanes it found.
template <int NW, bool SIGMA>
__device__ void hub_segments(const Eng<NW> &e, int L, int mode) {
  const int lane = lane_id();
  const u64 *fc = frs(e, L);
  const u64 *fn = frs(e, L + 1);
  u64 dag = 0;
  int ovf = 0;
  while (true) {
    unsigned si = 0;
    if (lane == 0) si = atomicAdd(&e.ctrl->work[3], 1u);
    si = bcast(si);
    if (si >= (unsigned)e.nsegs) break;
    const int4 sg = e.segs[si];
    const int w = sg.x;
    u64 tgt[NW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      tgt[j] = mode ? (e.all[j] & ~e.visit[(size_t)w * NW + j]) : fn[(size_t)w * NW + j];
      any |= tgt[j] != 0;
    }
    if (!any) continue;
    u64 found[NW], acc[NW];
    pull<NW, SIGMA>(e, fc, sg.y, sg.z, tgt, found, acc, &ovf, &dag);
    if (mode && lane == 0) {
#pragma unroll
      for (int j = 0; j < NW; j++)
        if (found[j]) atomicOr(&e.hub_found[(size_t)sg.w * NW + j], found[j]);
    }
    if (SIGMA) {
      u64 *srow = e.svm + (size_t)w * (kWL * NW);
#pragma unroll
      for (int j = 0; j < NW; j++) {
        if (acc[j] != 0) {
          u64 old = atomicAdd(&srow[j * kWL + lane], acc[j]);
          ovf |= old + acc[j] < acc[j];
        }
      }
    }
  }
  if (ballot(ovf != 0) && lane == 0) atomicOr(&e.ctrl->overflow, 1);
  if (lane == 0 && dag) atomicAdd(&e.ctrl->dag, dag);
}

// Hub commit: new lanes of every hub (bottom-up: found by its segments, then
// pushed; top-down: already pushed by the expansion), distances, visit mask.
template <int NW>
__device__ void hub_commit(const Eng<NW> &e, int L, int mode) {
  const int lane = lane_id();
  u64 *fn = frs(e, L + 1);
  int *Qn = qptr(e, L + 1);
  const int waves = gridDim.x * (blockDim.x / kWave);
  u64 fe_acc = 0, ud_acc = 0;
  for (int h = blockIdx.x * (blockDim.x / kWave) + wave_id(); h < e.nhubs; h += waves) {
    const int w = e.hubs[h];
    u64 vo[NW], nw[NW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      vo[j] = e.visit[(size_t)w * NW + j];
      nw[j] = mode ? e.hub_found[(size_t)h * NW + j] : fn[(size_t)w * NW + j];
      any |= nw[j] != 0;
    }
    if (mode && lane == 0) {
#pragma unroll
      for (int j = 0; j < NW; j++) e.hub_found[(size_t)h * NW + j] = 0;
    }
    if (!any) continue;
    commit<NW, false>(e, L, w, vo, nw, nullptr);
    if (mode && lane == 0) {
#pragma unroll
      for (int j = 0; j < NW; j++) fn[(size_t)w * NW + j] = nw[j];
      unsigned pos = atomicAdd(&e.ctrl->qlen[slot_of(L + 1)], 1u);
      Qn[pos] = w;
      if (e.stack) {
#pragma unroll
        for (int j = 0; j < NW; j++) e.smask[((size_t)e.ep[L + 1] + pos) * NW + j] = nw[j];
      }
    }
    const u64 deg = (u64)(e.R[w + 1] - e.R[w]);
    fe_acc += deg;
    if (full_mask<NW>(e, vo, nw)) ud_acc += deg;
  }
  if (lane == 0) {
    if (fe_acc) atomicAdd(&e.ctrl->fr_edges[slot_of(L + 1)], fe_acc);
    if (ud_acc) atomicAdd(&e.ctrl->unvis_dec, ud_acc);
  }
}

}  // namespace gu
