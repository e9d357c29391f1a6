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
  const int *R;        // out-edges (CSR): top-down expansion, Brandes successors
  const int *C;
  const int *RP;       // in-edges (the transpose; the same arrays when the
  const int *CP;       //   graph is undirected): pulls from the frontier
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
  u64 *acct;           // [3][kAcctMax][4] per-workgroup level accounting
  int grid_k[3];       // grid sizes of the three accounting kernels
  u64 all[NW];         // mask of active lanes per word
  int simple;          // adjacency rows strictly increasing: no parallel
                       //   edges, so level-1 vertices have one path per lane
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

// Per-level accounting of one wave: out-degree sum of the vertices it
// pushed (next frontier edges), of the vertices it completed (every lane
// reached), shortest-path DAG edges pulled, path-count overflow, and the
// (vertex, lane) pairs it committed, plain and weighted by out-degree (the
// reached count and TEPS units).  Summed per workgroup into a slot of
// e.acct (no shared counter), then by k_control.
struct Acc {
  u64 fe = 0, ud = 0, dag = 0, tp = 0, rc = 0;
  int ovf = 0;
};

constexpr int kAcctW = 6;  // words per accounting slot

// Lanes newly committed for a vertex of out-degree deg.
template <int NW>
__device__ __forceinline__ void acct_commit(Acc &a, const u64 *nw, u64 deg) {
  u64 pc = 0;
#pragma unroll
  for (int j = 0; j < NW; j++) pc += (u64)__popcll(nw[j]);
  a.rc += pc;
  a.tp += pc * deg;
}

constexpr int kAcctMax = 8192;  // largest grid of an accounting kernel

// Whole workgroup: every thread must call it once.
template <int NW>
__device__ void write_acct(const Eng<NW> &e, int kid, const Acc &a) {
  __shared__ u64 s_acc[16][kAcctW];
  const int lane = lane_id(), wid = wave_id();
  const bool ovf = ballot(a.ovf != 0) != 0;
  if (lane == 0) {
    s_acc[wid][0] = a.fe;
    s_acc[wid][1] = a.ud;
    s_acc[wid][2] = a.dag;
    s_acc[wid][3] = ovf ? 1ull : 0ull;
    s_acc[wid][4] = a.tp;
    s_acc[wid][5] = a.rc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t[kAcctW] = {};
    for (int w = 0; w < (int)blockDim.x / kWave; w++)
      for (int i = 0; i < kAcctW; i++) t[i] = i == 3 ? (t[i] | s_acc[w][i]) : t[i] + s_acc[w][i];
    u64 *slot = e.acct + ((size_t)kid * kAcctMax + blockIdx.x) * kAcctW;
    for (int i = 0; i < kAcctW; i++) slot[i] = t[i];
  }
}

// Per-wave queue buffer in LDS: pushes are appended with ballot ranks and
// written out, coalesced, behind one global atomic per flush.
constexpr int kPB = 128;

struct Pusher {
  int *buf;  // this wave's kPB LDS entries
  int cnt;   // wave-uniform fill
};

template <int NW>
__device__ __forceinline__ void pb_flush(const Eng<NW> &e, int L, Pusher &pb) {
  if (pb.cnt == 0) return;
  unsigned base = 0;
  if (lane_id() == 0) base = atomicAdd(&e.ctrl->qlen[slot_of(L + 1)], (unsigned)pb.cnt);
  base = (unsigned)__builtin_amdgcn_readfirstlane((int)base);
  int *Qn = qptr(e, L + 1);
  __builtin_amdgcn_wave_barrier();
  for (int i = lane_id(); i < pb.cnt; i += kWave) Qn[base + i] = pb.buf[i];
  __builtin_amdgcn_wave_barrier();
  pb.cnt = 0;
}

template <int NW>
__device__ __forceinline__ void pb_push(const Eng<NW> &e, int L, Pusher &pb, bool p, int w) {
  uint64_t m = ballot(p);
  if (!m) return;
  if (p) pb.buf[pb.cnt + rank_in(m)] = w;
  pb.cnt += __popcll(m);
  __builtin_amdgcn_wave_barrier();
  if (pb.cnt > kPB - kWave) pb_flush<NW>(e, L, pb);
}

template <int NW>
__device__ __forceinline__ void load_mask(const u64 *p, u64 *f) {
#pragma unroll
  for (int j = 0; j < NW; j++) f[j] = p[j];
}

// Strip-mine the adjacency list [r0,r1) of a frontier vertex with masks f,
// 64 entries per step (uniform trip count across the wave).
template <int NW>
__device__ __forceinline__ void strip_wave(const Eng<NW> &e, int L, Pusher &pb, const u64 *f, int r0,
                                           int r1) {
  const int lane = lane_id();
  for (int e0 = r0; e0 < r1; e0 += kWave) {
    int ed = e0 + lane;
    bool valid = ed < r1;
    int w = valid ? e.C[ed] : 0;
    bool p = valid && discover<NW>(e, L, f, w);
    pb_push<NW>(e, L, pb, p, w);
  }
}

// ---- top-down strategies ------------------------------------------------

// shuffle_based.cuh:192-223: one wave per frontier vertex, waves fetch the
// next queue entries dynamically; lanes stride the adjacency list.
template <int NW>
__device__ void td_shuffle(const Eng<NW> &e, int L, Pusher &pb) {
  const int lane = lane_id();
  const int *Q = qptr(e, L);
  const unsigned qn = e.ctrl->qlen[slot_of(L)];
  const u64 *fc = frs(e, L);
  constexpr unsigned CH = 16;
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
      strip_wave<NW>(e, L, pb, f, e.R[v], e.R[v + 1]);
    }
  }
}

// warp_based.cu:89-166: every lane holds one frontier vertex; the wave
// repeatedly elects the highest lane that still has adjacency left
// (race_and_resolve) and strip-mines that lane's list.
template <int NW>
__device__ void td_warp_core(const Eng<NW> &e, int L, Pusher &pb, int v, int r, int re) {
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
    strip_wave<NW>(e, L, pb, f, rw, rew);
  }
}

template <int NW>
__device__ void td_warp(const Eng<NW> &e, int L, Pusher &pb) {
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
    td_warp_core<NW>(e, L, pb, v, r, re);
  }
}

// CTA_warp.cu:105-186 then :188-265: vertices whose degree exceeds the
// threshold are expanded by the whole workgroup, one at a time; the rest by
// the wave election above.
template <int NW>
__device__ void td_cta_warp(const Eng<NW> &e, int L, Pusher &pb) {
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
        pb_push<NW>(e, L, pb, p, w);
      }
      __syncthreads();
    }
    td_warp_core<NW>(e, L, pb, v, r, re);
  }
}

// scan_based.cu:105-169: a workgroup scans the degrees of its frontier
// tile, then gathers the tile's edges blockDim at a time.  Each thread finds
// the owner of its edge by binary search over the tile's scanned degrees in
// LDS, so a hub in the tile costs its edges / blockDim rounds, not a serial
// fill by the hub's own thread (2x slower BC on R-MAT s22 at 64 lanes).
template <int NW>
__device__ void td_scan(const Eng<NW> &e, int L, Pusher &pb) {
  __shared__ int s_off[1024];
  __shared__ int s_r[1024];
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
    int total;
    int rsv = block_exclusive_sum(re - r, s_scan, &total);
    s_vert[tid] = v;
    s_r[tid] = r;
    s_off[tid] = rsv;
    __syncthreads();
    for (int progress = 0; progress < total; progress += T) {
      const int cnt = total - progress < T ? total - progress : T;
      bool p = false;
      int w = 0;
      if (tid < cnt) {
        const int g = progress + tid;
        int a = 0, b = T;  // last slot with s_off <= g
        while (b - a > 1) {
          const int mid = (a + b) >> 1;
          if (s_off[mid] <= g) a = mid; else b = mid;
        }
        u64 f[NW];
        load_mask<NW>(fc + (size_t)s_vert[a] * NW, f);
        w = e.C[s_r[a] + (g - s_off[a])];
        p = discover<NW>(e, L, f, w);
      }
      pb_push<NW>(e, L, pb, p, w);
    }
    __syncthreads();
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
__device__ void td_edge_parallel(const Eng<NW> &e, int L, Pusher &pb) {
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
      pb_push<NW>(e, L, pb, p, w);
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
    int v = valid ? e.CP[ed] : 0;
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

// Group pull with path counts.  A wave takes up to 64 vertices (lane i owns
// vertex w_i with target lanes tgt_i and in-edge range [r0_i, r0_i + deg_i);
// deg_i = 0 drops the vertex).  The owned in-edge ranges are concatenated by
// a wave prefix sum and walked 64 entries per step, each lane finding its
// owner by binary search in LDS (load-balanced search), so low-degree
// vertices share a step and vertices with nothing left to find cost nothing.
// Predecessor rows are added in entry order, four loads in flight; when the
// owner changes the finished vertex is committed.  BU: found lanes are OR-ed
// per owner in LDS and the discovered vertices are pushed; TD: the target
// lanes are the lanes pushed this level.
template <int NW>
struct GroupLds {
  u64 tgt[kWave][NW];
  u64 found[kWave][NW];
  // visit mask of an owner vertex, loaded with the group (up to two mask
  // words: for four the larger LDS cost 5% of occupancy-bound throughput)
  u64 vis[kWave][NW <= 2 ? NW : 1];
  int off[kWave + 1];
  int w[kWave];
  int r0[kWave];
  int aux[kWave];   // hub index of a segment owner; out-degree of a vertex owner
};

// What a group pull does with a finished owner.
enum PullMode {
  PM_TD = 0,   // top-down finalize: commit the lanes pushed this level
  PM_BU = 1,   // bottom-up: commit the lanes found, push the vertex
  PM_SEG = 2,  // hub segment: atomically add partial path counts (and, when
               // bottom-up, OR the found lanes into the hub's mask)
};

__device__ __forceinline__ int owner_of(const int *off, int t) {
  int a = 0, b = kWave;  // largest i with off[i] <= t
#pragma unroll
  for (int it = 0; it < 6; it++) {
    int mid = (a + b) >> 1;
    if (off[mid] <= t) a = mid; else b = mid;
  }
  return a;
}

// A vertex owner's out-degree (aux_l) and, for up to two mask words, its
// visit mask (vis_l) are loaded by the caller for all 64 owners at once:
// only the owner's own commit writes the mask during the level, so the
// commits need no dependent load of their own (+1.9% per s20 pass).
template <int NW, int MODE>
__device__ void pull_group(const Eng<NW> &e, int L, int w_l, const u64 *tgt_l, int r0_l, int deg_l,
                           GroupLds<NW> &g, Pusher &pb, Acc &a, int aux_l = 0, bool seg_bu = false,
                           const u64 *vis_l = nullptr) {
  constexpr bool BU = MODE == PM_BU;
  constexpr bool OR_FOUND = MODE != PM_TD;
  // predecessor rows in flight, measured on s20 BFS: 64 lanes 12 (8: -0.5%,
  // 16: equal at 80 VGPRs), 128 lanes 8 (4: -2%, 12: -4%); 256 lanes 6:
  // with 8 the expansion kernel spilled registers, and 1024-source
  // Brandes on s22 at width 256 took 409.8 ms against 401.0 with 6 (4:
  // 402.1)
  constexpr int D = NW == 1 ? 12 : (NW == 2 ? 8 : 6);
  const int lane = lane_id();
  const int inc = wave_inclusive_sum(deg_l);
  const int T = wave_total(inc);
  if (T == 0) return;
  g.off[lane] = inc - deg_l;
  if (lane == kWave - 1) g.off[kWave] = T;
  g.w[lane] = w_l;
  g.r0[lane] = r0_l;
  g.aux[lane] = aux_l;
#pragma unroll
  for (int j = 0; j < NW; j++) {
    g.tgt[lane][j] = tgt_l[j];
    g.found[lane][j] = 0;
    if constexpr (MODE != PM_SEG && NW <= 2) g.vis[lane][j] = vis_l[j];
  }
  __builtin_amdgcn_wave_barrier();
  const u64 *fc = frs(e, L);
  u64 *fn = frs(e, L + 1);
  int cur = -1;
  uint64_t disc = 0;
  u64 acc[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) acc[j] = 0;

  auto flush = [&](int slot) {
    const int w = g.w[slot];
    if (MODE == PM_SEG) {
      u64 *srow = e.svm + (size_t)w * (kWL * NW);
#pragma unroll
      for (int j = 0; j < NW; j++) {
        if (acc[j] != 0) {
          u64 old = atomicAdd(&srow[j * kWL + lane], acc[j]);
          a.ovf |= old + acc[j] < acc[j];
        }
      }
      if (seg_bu && lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++)
          if (g.found[slot][j]) atomicOr(&e.hub_found[(size_t)g.aux[slot] * NW + j], g.found[slot][j]);
      }
      return;
    }
    u64 vo[NW], nw[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) {
      nw[j] = BU ? g.found[slot][j] : g.tgt[slot][j];
      if constexpr (NW <= 2)
        vo[j] = g.vis[slot][j];
      else
        vo[j] = e.visit[(size_t)w * NW + j];
    }
    commit<NW, true>(e, L, w, vo, nw, acc);
    if (BU && lane == 0) {
#pragma unroll
      for (int j = 0; j < NW; j++) fn[(size_t)w * NW + j] = nw[j];
    }
    const u64 deg = (u64)g.aux[slot];
    a.fe += deg;
    if (full_mask<NW>(e, vo, nw)) a.ud += deg;
    acct_commit<NW>(a, nw, deg);
    disc |= 1ull << slot;
  };

  // A step is 64 in-edges of the group.  The next step's in-edge is loaded
  // before the row loads of the step in hand; with one mask word (64
  // lanes) the next step's frontier masks are too, and the in-edge two
  // steps ahead, so a step waits only for its rows (+0.9% per s20 pass).
  // Wider masks keep the one-step form: the extra registers would cost
  // them occupancy.
  constexpr bool AHEAD = NW == 1;
  int t = lane;
  bool valid = t < T;
  int own = valid ? owner_of(g.off, t) : 0;
  int v = valid ? e.CP[g.r0[own] + t - g.off[own]] : 0;
  u64 fcv[NW];
  int tn = 0, ownn = 0, vn = 0;
  bool validn = false;
  if constexpr (AHEAD) {
#pragma unroll
    for (int j = 0; j < NW; j++) fcv[j] = valid ? fc[(size_t)v * NW + j] : 0ull;
    tn = t + kWave;
    validn = tn < T;
    ownn = validn ? owner_of(g.off, tn) : 0;
    vn = validn ? e.CP[g.r0[ownn] + tn - g.off[ownn]] : 0;
  }
  for (int t0 = 0; t0 < T; t0 += kWave) {
    u64 m[NW];
    bool has = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      if constexpr (AHEAD)
        m[j] = fcv[j] & g.tgt[own][j];
      else
        m[j] = valid ? (fc[(size_t)v * NW + j] & g.tgt[own][j]) : 0ull;
      has |= m[j] != 0;
    }
    int tnn = 0, ownnn = 0, vnn = 0;
    bool validnn = false;
    if constexpr (AHEAD) {
#pragma unroll
      for (int j = 0; j < NW; j++) fcv[j] = validn ? fc[(size_t)vn * NW + j] : 0ull;
      tnn = tn + kWave;
      validnn = tnn < T;
      ownnn = validnn ? owner_of(g.off, tnn) : 0;
      vnn = validnn ? e.CP[g.r0[ownnn] + tnn - g.off[ownnn]] : 0;
    } else {
      tn = t + kWave;
      validn = tn < T;
      ownn = validn ? owner_of(g.off, tn) : 0;
      vn = validn ? e.CP[g.r0[ownn] + tn - g.off[ownn]] : 0;
    }
    if (OR_FOUND) {
#pragma unroll
      for (int j = 0; j < NW; j++)
        if (m[j]) atomicOr(&g.found[own][j], m[j]);
      __builtin_amdgcn_wave_barrier();
    }
    uint64_t pm = ballot(has);
    while (pm) {
      int k[D];
#pragma unroll
      for (int u = 0; u < D; u++) {
        if (pm) {
          k[u] = __ffsll((unsigned long long)pm) - 1;
          pm &= pm - 1;
        } else {
          k[u] = -1;
        }
      }
      u64 x[D][NW];
      int ow[D];
#pragma unroll
      for (int u = 0; u < D; u++) {
        ow[u] = -1;
        if (k[u] >= 0) {
          ow[u] = readlane(own, k[u]);
          const int vk = readlane(v, k[u]);
          const u64 *row = e.svm + (size_t)vk * (kWL * NW);
#pragma unroll
          for (int j = 0; j < NW; j++) {
            const u64 mk = readlane64(m[j], k[u]);
            a.dag += (u64)__popcll(mk);
            x[u][j] = ((mk >> lane) & 1ull) ? row[j * kWL + lane] : 0ull;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < D; u++) {
        if (ow[u] < 0) continue;
        if (ow[u] != cur) {
          if (cur >= 0) flush(cur);
          cur = ow[u];
#pragma unroll
          for (int j = 0; j < NW; j++) acc[j] = 0;
        }
#pragma unroll
        for (int j = 0; j < NW; j++) {
          u64 sum = acc[j] + x[u][j];
          a.ovf |= sum < x[u][j];
          acc[j] = sum;
        }
      }
    }
    t = tn;
    valid = validn;
    own = ownn;
    v = vn;
    if constexpr (AHEAD) {
      tn = tnn;
      validn = validnn;
      ownn = ownnn;
      vn = vnn;
    }
  }
  if (cur >= 0) flush(cur);
  if (BU) pb_push<NW>(e, L, pb, (disc >> lane) & 1ull, g.w[lane]);
  __builtin_amdgcn_wave_barrier();
}

// Dynamic LDS of the level kernels: one GroupLds and one push buffer per
// wave, sized by the launch.
extern __shared__ __align__(16) unsigned char gu_dyn_lds[];

template <int NW>
__host__ __device__ constexpr size_t level_lds_bytes(int waves) {
  return (size_t)waves * (sizeof(GroupLds<NW>) + kPB * sizeof(int));
}
template <int NW>
__device__ __forceinline__ GroupLds<NW> &wave_group() {
  return reinterpret_cast<GroupLds<NW> *>(gu_dyn_lds)[wave_id()];
}
template <int NW>
__device__ __forceinline__ int *wave_pushbuf() {
  const int waves = blockDim.x / kWave;
  return reinterpret_cast<int *>(gu_dyn_lds + (size_t)waves * sizeof(GroupLds<NW>)) + wave_id() * kPB;
}

// Bottom-up level: every vertex that some lane has not reached pulls from
// its in-neighbours on the frontier.  Waves take precomputed blocks of at
// most 64 consecutive vertices with a bounded in-edge count (hubs are left
// to the segment pass), statically: the blocks are already edge-balanced,
// and waves running side by side work on neighbouring vertex ranges.
// (Claiming blocks from a level counter instead measured 15% slower.)
template <int NW, bool SIGMA>
__device__ void bu_level(const Eng<NW> &e, int L, Pusher &pb, GroupLds<NW> &g, Acc &a) {
  const int lane = lane_id();
  const u64 *fc = frs(e, L);
  u64 *fn = frs(e, L + 1);
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int bi = blockIdx.x * (blockDim.x / kWave) + wave_id(); bi < e.nblocks; bi += waves) {
    const int2 blk = e.blocks[bi];
    const int nblk = blk.y - blk.x;
    if (SIGMA) {
      const bool have = lane < nblk;
      const int w = blk.x + (have ? lane : 0);
      // every lane loads (w is a vertex of the block either way) and the
      // values are selected afterwards: under branches the row offsets
      // waited for the visit masks
      u64 vraw[NW];
#pragma unroll
      for (int j = 0; j < NW; j++) vraw[j] = e.visit[(size_t)w * NW + j];
      const int rp0 = e.RP[w], rp1 = e.RP[w + 1], ro0 = e.R[w], ro1 = e.R[w + 1];
      u64 tgt[NW], vis[NW];
      bool any = false;
#pragma unroll
      for (int j = 0; j < NW; j++) {
        vis[j] = have ? vraw[j] : 0ull;
        tgt[j] = have ? (e.all[j] & ~vis[j]) : 0ull;
        any |= tgt[j] != 0;
      }
      const int r0 = have ? rp0 : 0;
      const int deg = any ? rp1 - r0 : 0;
      const int odeg = any ? ro1 - ro0 : 0;
      pull_group<NW, PM_BU>(e, L, w, tgt, r0, deg, g, pb, a, odeg, false, vis);
      continue;
    }
    // distances only: one vertex at a time, stopping as soon as every
    // target lane has been found
    uint64_t disc = 0;
    for (int i = 0; i < nblk; i++) {
      const int w = blk.x + i;
      const int r0 = e.RP[w], r1 = e.RP[w + 1];
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
      pull<NW, false>(e, fc, r0, r1, tgt, found, acc, &a.ovf, &a.dag);
      bool anyn = false;
#pragma unroll
      for (int j = 0; j < NW; j++) anyn |= found[j] != 0;
      if (!anyn) continue;
      commit<NW, false>(e, L, w, vo, found, acc);
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++) fn[(size_t)w * NW + j] = found[j];
      }
      disc |= 1ull << i;
      const u64 deg = (u64)(e.R[w + 1] - e.R[w]);
      a.fe += deg;
      if (full_mask<NW>(e, vo, found)) a.ud += deg;
      acct_commit<NW>(a, found, deg);
    }
    pb_push<NW>(e, L, pb, (disc >> lane) & 1ull, blk.x + lane);
  }
}

// Top-down finalize: for every non-hub vertex pushed this level, commit its
// new lanes and (with path counts) pull sigma from its frontier
// in-neighbours.  Waves take 32-entry chunks of the queue in a fixed
// interleave; hubs are committed by the hub pass.
template <int NW, bool SIGMA>
__device__ void td_finalize(const Eng<NW> &e, int L, GroupLds<NW> &g, Acc &a) {
  constexpr int CH = 32;
  const int lane = lane_id();
  const int *Qn = qptr(e, L + 1);
  const unsigned qn = e.ctrl->qlen[slot_of(L + 1)];
  const u64 *fn = frs(e, L + 1);
  const unsigned waves = gridDim.x * (blockDim.x / kWave);
  Pusher none{nullptr, 0};
  for (unsigned base = (blockIdx.x * (blockDim.x / kWave) + wave_id()) * CH; base < qn; base += waves * CH) {
    const int cnt = qn - base < (unsigned)CH ? (int)(qn - base) : CH;
    const bool have = lane < cnt;
    const int w = have ? Qn[base + lane] : 0;
    const int r0 = have ? e.RP[w] : 0;
    const int deg = have ? e.RP[w + 1] - r0 : 0;
    const bool hub = deg > e.hub_deg;  // committed by the hub pass
    u64 nw[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) nw[j] = (have && !hub) ? fn[(size_t)w * NW + j] : 0ull;
    // a level-1 vertex of a graph without parallel edges has exactly one
    // path (its edge from the source) in each lane that reached it: no pull
    const bool unit = SIGMA && L == 0 && e.simple;
    if (SIGMA && !unit) {
      u64 vis[NW];
#pragma unroll
      for (int j = 0; j < NW; j++) vis[j] = (have && !hub) ? e.visit[(size_t)w * NW + j] : 0ull;
      const int odeg = (have && !hub) ? e.R[w + 1] - e.R[w] : 0;
      pull_group<NW, PM_TD>(e, L, w, nw, r0, hub ? 0 : deg, g, none, a, odeg, false, vis);
      continue;
    }
    u64 one[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) one[j] = 1;
    // every entry's visit mask and out-degree are loaded at once, lane by
    // lane; the commits below then only store
    const bool mine = have && !hub;
    u64 vis[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) vis[j] = mine ? e.visit[(size_t)w * NW + j] : 0ull;
    const int odeg = mine ? e.R[w + 1] - e.R[w] : 0;
    uint64_t todo = ballot(mine);
    while (todo) {
      const int k = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int wk = readlane(w, k);
      u64 vo[NW], nk[NW];
#pragma unroll
      for (int j = 0; j < NW; j++) {
        vo[j] = readlane64(vis[j], k);
        nk[j] = readlane64(nw[j], k);
      }
      commit<NW, SIGMA>(e, L, wk, vo, nk, one);
      if (SIGMA) {
#pragma unroll
        for (int j = 0; j < NW; j++) a.dag += (u64)__popcll(nk[j]);
      }
      const u64 d = (u64)readlane(odeg, k);
      a.fe += d;
      if (full_mask<NW>(e, vo, nk)) a.ud += d;
      acct_commit<NW>(a, nk, d);
    }
  }
}

// Hub segments (both directions): each wave pulls a pair of segments of hub
// adjacency lists for the hubs' target lanes (top-down: the lanes pushed this
// level; bottom-up: the lanes not yet reached) and adds its partial path
// counts with 64-bit atomics -- exact, so the sum does not depend on the
// order of the segments.  Bottom-up also ORs the lanes it found.
template <int NW, bool SIGMA>
__device__ void hub_segments(const Eng<NW> &e, int L, int mode, GroupLds<NW> &g, Acc &a) {
  constexpr int G = 2;  // segments per wave step (1 and 4 measured no better)
  const int lane = lane_id();
  const u64 *fc = frs(e, L);
  const u64 *fn = frs(e, L + 1);
  const int waves = gridDim.x * (blockDim.x / kWave);
  Pusher none{nullptr, 0};
  for (int s0 = (blockIdx.x * (blockDim.x / kWave) + wave_id()) * G; s0 < e.nsegs; s0 += waves * G) {
    if (SIGMA) {
      const bool have = lane < G && s0 + lane < e.nsegs;
      const int4 sg = have ? e.segs[s0 + lane] : make_int4(0, 0, 0, 0);
      u64 tgt[NW];
      bool any = false;
#pragma unroll
      for (int j = 0; j < NW; j++) {
        tgt[j] = !have ? 0ull
                       : (mode ? (e.all[j] & ~e.visit[(size_t)sg.x * NW + j]) : fn[(size_t)sg.x * NW + j]);
        any |= tgt[j] != 0;
      }
      pull_group<NW, PM_SEG>(e, L, sg.x, tgt, sg.y, any ? sg.z - sg.y : 0, g, none, a, sg.w, mode != 0);
      continue;
    }
    for (int si = s0; si < s0 + G && si < e.nsegs; si++) {
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
      pull<NW, false>(e, fc, sg.y, sg.z, tgt, found, acc, &a.ovf, &a.dag);
      if (mode && lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++)
          if (found[j]) atomicOr(&e.hub_found[(size_t)sg.w * NW + j], found[j]);
      }
    }
  }
}

// Top-down level 0 with path counts.  The frontier is the sources, each
// with path count 1 in its own lanes, so a hub reached at level 1 has, in
// lane b, one path per edge from b's source to it.  Rather than pulling
// every newly reached hub's whole in-edge list (the segments), the sources'
// out-edges are walked once, in 64-edge chunks spread over all waves, and
// each edge into a hub adds 1 to the hub's row in the lanes it reached at
// level 1: exact u64 atomics into the zeroed hub rows over the same edge
// multiset as the pull, so the same sums.
template <int NW>
__device__ void hub_push_sources(const Eng<NW> &e, Acc &a) {
  const int lane = lane_id();
  const int *Q = qptr(e, 0);
  const int qn = (int)e.ctrl->qlen[0];  // distinct sources: at most 64 * NW
  const u64 *fc = frs(e, 0);
  const u64 *fn = frs(e, 1);
  const int waves = gridDim.x * (blockDim.x / kWave);
  const int gw = blockIdx.x * (blockDim.x / kWave) + wave_id();
  int uq[NW], r0[NW], r1[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const int i = j * kWave + lane;
    uq[j] = i < qn ? Q[i] : 0;
    r0[j] = i < qn ? e.R[uq[j]] : 0;
    r1[j] = i < qn ? e.R[uq[j] + 1] : 0;
  }
  long long c = 0;  // chunks of the entries before this one
#pragma unroll
  for (int j = 0; j < NW; j++) {
    for (int l = 0; l < kWave && j * kWave + l < qn; l++) {
      const int s0 = readlane(r0[j], l), s1 = readlane(r1[j], l);
      const int nch = (s1 - s0 + kWave - 1) / kWave;
      long long k = ((long long)gw - c) % waves;  // this wave's first chunk of the entry
      if (k < 0) k += waves;
      c += nch;
      if (k >= nch) continue;
      const int u = readlane(uq[j], l);
      u64 f[NW];
#pragma unroll
      for (int jj = 0; jj < NW; jj++) f[jj] = fc[(size_t)u * NW + jj];
      for (; k < nch; k += waves) {
        const int t = s0 + (int)k * kWave + lane;
        const bool ok = t < s1;
        const int w = ok ? e.C[t] : 0;
        const bool hub = ok && e.RP[w + 1] - e.RP[w] > e.hub_deg;
        u64 m[NW];
        bool any = false;
#pragma unroll
        for (int jj = 0; jj < NW; jj++) {
          m[jj] = hub ? (fn[(size_t)w * NW + jj] & f[jj]) : 0ull;
          any |= m[jj] != 0;
        }
        uint64_t pm = ballot(any);
        while (pm) {
          const int kk = __ffsll((unsigned long long)pm) - 1;
          pm &= pm - 1;
          u64 *row = e.svm + (size_t)readlane(w, kk) * (kWL * NW);
#pragma unroll
          for (int jj = 0; jj < NW; jj++) {
            const u64 mk = readlane64(m[jj], kk);
            a.dag += (u64)__popcll(mk);
            if ((mk >> lane) & 1ull) atomicAdd(&row[jj * kWL + lane], 1ull);
          }
        }
      }
    }
  }
}

// Hub commit, one lane per hub: new lanes (bottom-up: found by the hub's
// segments, then pushed; top-down: already pushed by the expansion),
// visit mask and distances.
template <int NW>
__device__ void hub_commit(const Eng<NW> &e, int L, int mode, Pusher &pb, Acc &a) {
  const int lane = lane_id();
  u64 *fn = frs(e, L + 1);
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int h0 = (blockIdx.x * (blockDim.x / kWave) + wave_id()) * kWave; h0 < e.nhubs; h0 += waves * kWave) {
    const int h = h0 + lane;
    const bool have = h < e.nhubs;
    const int w = have ? e.hubs[h] : 0;
    u64 vo[NW], nw[NW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      vo[j] = have ? e.visit[(size_t)w * NW + j] : 0ull;
      nw[j] = !have ? 0ull : (mode ? e.hub_found[(size_t)h * NW + j] : fn[(size_t)w * NW + j]);
      any |= nw[j] != 0;
      if (have && mode) e.hub_found[(size_t)h * NW + j] = 0;
    }
    if (any) {
#pragma unroll
      for (int j = 0; j < NW; j++) {
        e.visit[(size_t)w * NW + j] = vo[j] | nw[j];
        if (mode) fn[(size_t)w * NW + j] = nw[j];
      }
    }
    // distances: every lane writes its own entry of each committed hub row
    uint64_t pm = ballot(any);
    while (pm) {
      const int k = __ffsll((unsigned long long)pm) - 1;
      pm &= pm - 1;
      const int wk = readlane(w, k);
      int *drow = e.dvm + (size_t)wk * (kWL * NW);
#pragma unroll
      for (int j = 0; j < NW; j++) {
        const u64 nk = readlane64(nw[j], k);
        if ((nk >> lane) & 1ull) drow[j * kWL + lane] = L + 1;
      }
    }
    if (mode) pb_push<NW>(e, L, pb, any, w);
    u64 fe = any ? (u64)(e.R[w + 1] - e.R[w]) : 0ull;
    u64 ud = (any && full_mask<NW>(e, vo, nw)) ? fe : 0ull;
    Acc hc;  // this lane's hub
    if (any) acct_commit<NW>(hc, nw, fe);
    u64 tp = hc.tp, rc = hc.rc;
    for (int o = 32; o >= 1; o >>= 1) {
      fe += __shfl_xor(fe, o, kWave);
      ud += __shfl_xor(ud, o, kWave);
      tp += __shfl_xor(tp, o, kWave);
      rc += __shfl_xor(rc, o, kWave);
    }
    a.fe += fe;
    a.ud += ud;
    a.tp += tp;
    a.rc += rc;
  }
}

}  // namespace gu
