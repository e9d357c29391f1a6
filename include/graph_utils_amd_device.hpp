/*
 * graph_utils_amd_device.hpp -- device-side API of the MI355X (gfx950)
 * multi-source search path, for HIP kernels.
 *
 * The reference exposes a device API next to its host calls: the warp
 * helpers of kernel_utils.cuh, race_and_resolve_warp (race_and_resolve.cuh),
 * the load-balanced search device functions and kernels
 * (load_balanced_search.cuh:10-13) and the hook-driven multi_search engine
 * (multi_search/shuffle_based.cuh:24-286).  This header gives each of them a
 * 64-lane form with the same arguments and meaning:
 *
 *   lane_id, wave_id, get_bit, set_bit        kernel_utils.cuh:8-35
 *   race_and_resolve_wave                     race_and_resolve.cuh:13-25
 *   atomic_agg_inc                            shuffle_based.cuh:58-76
 *   load_balance_search_wave / _block         load_balanced_search.cu:63-192
 *   extract_edges_wave / _block (kernels)     load_balanced_search.cu:194-202
 *   pitch, get_row, outdegree                 shuffle_based.cuh:24-54
 *   multi_search<F1..F7> (hook engine)        shuffle_based.cuh:80-286
 *
 * Differences that follow from the hardware:
 *  - A wavefront is 64 lanes; ballots are 64-bit and lane ranks come from
 *    mbcnt, so none of the 32-bit lane masks of the reference appear.
 *  - Scans use DPP row shifts and row broadcasts inside a wave and LDS
 *    across the waves of a workgroup (the reference uses CUB).
 *  - load_balance_search_wave synchronises only its own wave, so several
 *    waves of one workgroup may run it at once on different data (the
 *    reference's warp form calls __syncthreads inside a warp routine).
 *  - The load-balanced search locates each item's segment by a galloping
 *    search that starts at the segment of the thread's previous item, where
 *    the reference walks the scan one entry at a time; the output is the
 *    same: result[i] = (last segment j with scanned_edges[j] <= i).
 *
 * Compile with hipcc --offload-arch=gfx950.  Everything is header-only; the
 * kernels are templates whose parameters all have defaults, so
 * extract_edges_block<<<1, 512>>>(...) launches them as in lbs.cu:70.
 */
#ifndef GRAPH_UTILS_AMD_DEVICE_HPP
#define GRAPH_UTILS_AMD_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <climits>
#include <stddef.h>
#include <stdint.h>

namespace gu {

constexpr int kWave = 64;

/* ---- kernel_utils (kernel_utils.cuh:8-35) ------------------------------ */

/* Lane of this thread in its wavefront, 0..63. */
__device__ __forceinline__ int lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
/* Wavefront of this thread in its workgroup (getWarpId: the flattened thread
 * index over the wave size). */
__device__ __forceinline__ int wave_id() {
  return (int)((threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x) / kWave;
}
/* Bit `pos` of `val` (getBit, a bit-field extract of width 1). */
__device__ __forceinline__ int get_bit(int val, int pos) {
  return (int)__builtin_amdgcn_ubfe((uint32_t)val, (uint32_t)pos, 1u);
}
/* `val` with bit `pos` replaced by the low bit of `to_insert` (setBit, a
 * bit-field insert of width 1). */
__device__ __forceinline__ unsigned set_bit(unsigned val, unsigned to_insert, int pos) {
  return (val & ~(1u << pos)) | ((to_insert & 1u) << pos);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

/* Number of set bits of `mask` in the lanes below this one. */
__device__ __forceinline__ int rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

/* ---- race_and_resolve (race_and_resolve.cuh:13-25) ---------------------- */

/* Every lane of the wave calls this.  Returns, in every lane, the highest
 * lane whose predicate is nonzero; when no lane's predicate holds each lane
 * gets its own lane id. */
__device__ __forceinline__ int race_and_resolve_wave(int predicate) {
  const uint64_t vote = ballot(predicate != 0);
  if (vote == 0) return lane_id();
  return 63 - __clzll((long long)vote);
}

/* ---- wave-aggregated increment (shuffle_based.cuh:58-76) ---------------- */

/* atomicAggInc: each active lane of the wave takes a distinct slot
 * ctr_old + (its rank among the active lanes); one atomic per wave, issued
 * by the lowest active lane.  `ctr` may point to global or LDS memory. */
__device__ __forceinline__ int atomic_agg_inc(int *ctr) {
  const uint64_t m = ballot(true);
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(ctr, __popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  return base + rank_in(m);
}

/* The same with a predicate instead of control flow: lanes with p true take
 * slots; the others get -1.  Every lane of the wave calls it. */
__device__ __forceinline__ int atomic_agg_inc(unsigned int *ctr, bool p) {
  const uint64_t m = ballot(p);
  if (m == 0) return -1;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned int base = 0;
  if (lane_id() == leader) base = atomicAdd(ctr, (unsigned int)__popcll(m));
  base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);
  return p ? (int)(base + rank_in(m)) : -1;
}

/* ---- scans ------------------------------------------------------------- */

/* DPP move: row_shr:k is 0x110+k, row_bcast:15 is 0x142, row_bcast:31 is
 * 0x143. */
template <int CTRL, int ROWMASK, bool BOUND0>
__device__ __forceinline__ int dpp_i32(int old, int src) {
  return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWMASK, 0xf, BOUND0);
}

/* Inclusive prefix sum over the 64 lanes. */
__device__ __forceinline__ int wave_inclusive_sum(int x) {
  // within each row of 16 lanes: shifts of 1, 2, 4, 8 (lanes shifted in from
  // outside the row read 0)
  x += dpp_i32<0x111, 0xf, true>(0, x);
  x += dpp_i32<0x112, 0xf, true>(0, x);
  x += dpp_i32<0x114, 0xf, true>(0, x);
  x += dpp_i32<0x118, 0xf, true>(0, x);
  // row totals: lane 15 of row r into row r+1 (rows 1 and 3), then lane 31
  // into rows 2 and 3
  x += dpp_i32<0x142, 0xa, false>(0, x);
  x += dpp_i32<0x143, 0xc, false>(0, x);
  return x;
}

__device__ __forceinline__ int wave_total(int inclusive) { return __builtin_amdgcn_readlane(inclusive, kWave - 1); }

/* Exclusive workgroup scan (blockDim.x a multiple of 64, at most 1024).
 * `lds` needs blockDim.x/64 + 1 ints.  Returns the exclusive prefix; *total
 * receives the workgroup sum.  Contains __syncthreads(). */
__device__ __forceinline__ int block_exclusive_sum(int x, int *lds, int *total) {
  const int lane = lane_id();
  const int wid = (int)threadIdx.x / kWave;
  const int nw = (int)blockDim.x / kWave;
  const int inc = wave_inclusive_sum(x);
  if (lane == kWave - 1) lds[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    const int t = lane < nw ? lds[lane] : 0;
    const int ti = wave_inclusive_sum(t);
    if (lane < nw) lds[lane] = ti - t;  // exclusive wave offsets
    if (lane == kWave - 1) lds[nw] = ti;
  }
  __syncthreads();
  const int res = lds[wid] + inc - x;
  *total = lds[nw];
  __syncthreads();
  return res;
}

/* ---- load-balanced search (load_balanced_search.cuh:6-13) --------------- */

/* The reference's launch shape for the workgroup form: 512 threads, each
 * scanning 2 counts per step (load_balanced_search.cuh:7-8). */
constexpr int LBS_BLOCK_SIZE = 512;
constexpr int LBS_ITEMS_PER_THREAD = 2;

/* Largest j in [lo, S) with scan[j] <= i, given scan[lo] <= i: galloping
 * from lo, then a binary search. */
__device__ __forceinline__ int lbs_segment_from(const int *scan, int S, int lo, int i) {
  int step = 1, hi = lo + 1;
  while (hi < S && scan[hi] <= i) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > S) hi = S;
  // scan[lo] <= i; scan[hi] > i or hi == S
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (scan[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

/* One wavefront: scanned_edges[0..S) = exclusive prefix sum of
 * edge_counts[0..S), *edge_frontier_size = W = the total (written by lane
 * 0), result[i] = the last segment j with scanned_edges[j] <= i for every i
 * in [0, W) (zero-count segments own no item).  Counts must be >= 0, S >= 1
 * and W < 2^31.  Every lane of the wave calls it with the same arguments;
 * no workgroup barrier is used. */
__device__ __forceinline__ void load_balance_search_wave(const int vertex_frontier_size, int *edge_frontier_size,
                                                         const int *edge_counts, int *scanned_edges,
                                                         int *result) {
  const int S = vertex_frontier_size;
  const int lane = lane_id();
  int total = 0;
  for (int base = 0; base < S; base += kWave) {
    const int i = base + lane;
    const int c = i < S ? edge_counts[i] : 0;
    const int inc = wave_inclusive_sum(c);
    if (i < S) scanned_edges[i] = total + inc - c;
    total += wave_total(inc);
  }
  if (lane == 0) *edge_frontier_size = total;
  // the wave reads back the scan its own lanes wrote
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
  int seg = 0;
  for (int i = lane; i < total; i += kWave) {
    seg = lbs_segment_from(scanned_edges, S, seg, i);
    result[i] = seg;
  }
}

/* The workgroup form (blockDim.x a multiple of 64, at most 1024; the
 * reference uses LBS_BLOCK_SIZE): the same outputs as the wave form, the
 * counts scanned LBS_ITEMS_PER_THREAD per thread and step, then each thread
 * maps every blockDim.x-th item.  Every thread of the workgroup calls it. */
__device__ __forceinline__ void load_balance_search_block(const int vertex_frontier_size, int *edge_frontier_size,
                                                          const int *edge_counts, int *scanned_edges,
                                                          int *result) {
  __shared__ int s_scan[1024 / kWave + 1];
  const int S = vertex_frontier_size;
  const int T = (int)blockDim.x, tid = (int)threadIdx.x;
  constexpr int IT = LBS_ITEMS_PER_THREAD;
  int total = 0;
  for (int base = 0; base < S; base += T * IT) {
    int c[IT], sum = 0;
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int i = base + IT * tid + j;
      c[j] = i < S ? edge_counts[i] : 0;
      sum += c[j];
    }
    int step_total;
    int run = total + block_exclusive_sum(sum, s_scan, &step_total);
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int i = base + IT * tid + j;
      if (i < S) scanned_edges[i] = run;
      run += c[j];
    }
    total += step_total;
  }
  if (tid == 0) *edge_frontier_size = total;
  __syncthreads();  // the scan entries written above are visible to the workgroup
  int seg = 0;
  for (int i = tid; i < total; i += T) {
    seg = lbs_segment_from(scanned_edges, S, seg, i);
    result[i] = seg;
  }
}

/* extract_edges_wave / extract_edges_block (load_balanced_search.cu:194-202):
 * kernels running the two forms, *edges = W.  The wave form runs in the
 * first wave of each workgroup only (launch it with one wave). */
template <int UNUSED = 0>
__global__ void extract_edges_wave(int vertex_frontier_size, int *edge_counts, int *scanned_edges, int *result,
                                   int *edges) {
  if (threadIdx.x < (unsigned)kWave)
    load_balance_search_wave(vertex_frontier_size, edges, edge_counts, scanned_edges, result);
}

template <int UNUSED = 0>
__global__ __launch_bounds__(1024) void extract_edges_block(int vertex_frontier_size, int *edge_counts,
                                                            int *scanned_edges, int *result, int *edges) {
  load_balance_search_block(vertex_frontier_size, edges, edge_counts, scanned_edges, result);
}

/* ---- the hook-driven multi-search engine (shuffle_based.cuh:24-286) ------ */

/* Row pitches in bytes of the per-workgroup arrays (shuffle_based.cuh:24-41):
 * row b of array X starts at (char *)X + b * pitch.X. */
struct pitch {
  size_t d = 0, sigma = 0, delta = 0, bc = 0, Q = 0, Q2 = 0, S = 0, endpoints = 0, edge_counts = 0,
         scanned_edges = 0, LBS = 0, Q_small = 0, Q2_small = 0;
};

/* This workgroup's row of a pitched array (shuffle_based.cuh:44-48). */
template <class T>
__device__ __forceinline__ T *get_row(T *data, size_t p) {
  return (T *)((char *)data + blockIdx.x * p);
}

/* Out-degree of v (shuffle_based.cuh:51-54). */
__device__ __forceinline__ int outdegree(int v, const int *R) { return R[v + 1] - R[v]; }

/* Appends w to a queue whose length lives in LDS (the push of a visit
 * hook): one LDS atomic per wave for all its active lanes. */
__device__ __forceinline__ void ms_push(int *Q_row, int *Q_len, int w) { Q_row[atomic_agg_inc(Q_len)] = w; }

/* multi_search: each workgroup runs the sources start + blockIdx.x,
 * + gridDim.x, ... below end (one source at a time, the whole level loop in
 * this call).  For a source i:
 *   d_row (row blockIdx.x of d) gets 0 at i and INT_MAX elsewhere, with
 *   initLocal(k, i) called for every vertex k;
 *   thread 0 puts i on the queue and calls initStack(i);
 *   a source of out-degree 0 is skipped (no further hook is called);
 *   each level: every wave takes frontier entries v of Q one at a time
 *   (an LDS counter hands out the next one) and its lanes stride v's
 *   adjacency 64 at a time calling visitVertex(d_row, v, w, Q2_row, &Q2_len,
 *   Q2_small_row, &Q2_small_len); then each entry of Q_small is handled by
 *   one thread walking its whole adjacency (visits may push to either
 *   queue; the pushes use the LDS lengths passed in);
 *   when both next queues are empty: getMax(d_row[k]) for every vertex k and
 *   the level loop ends; otherwise insertStack(Q2_row, kk) for every new
 *   entry kk of Q2, then of Q2_small, the queues swap, and thread 0 calls
 *   updateEndpoints(Q2_len) with the new frontier's length;
 *   after the level loop, every thread calls
 *   dependencyAccum(d_row, i, threadIdx.x, lane).
 * The hooks run under the same barriers as in the reference, so hook code
 * written for it keeps its meaning with 64-lane waves.  blockDim.x must be a
 * multiple of 64. */
template <class F1, class F2, class F3, class F4, class F5, class F6, class F7>
__device__ void multi_search(const int *R, const int *C, const int n, int *d, int *Q, int *Q2, int *Q_small,
                             int *Q2_small, const pitch p, const int start, const int end, F1 getMax,
                             F2 initLocal, F3 visitVertex, F4 initStack, F5 insertStack, F6 updateEndpoints,
                             F7 dependencyAccum) {
  __shared__ int s_len[4];  // Q, Q2, Q_small, Q2_small lengths
  __shared__ int s_next;    // next frontier entry handed to a wave
  const int tid = (int)threadIdx.x, T = (int)blockDim.x;
  const int lane = lane_id(), wid = tid / kWave, nwaves = T / kWave;
  int *Q_row = get_row(Q, p.Q), *Q2_row = get_row(Q2, p.Q2);
  int *Qs_row = get_row(Q_small, p.Q_small), *Q2s_row = get_row(Q2_small, p.Q2_small);
  for (int i = (int)blockIdx.x + start; i < end; i += (int)gridDim.x) {
    int *d_row = get_row(d, p.d);
    for (int k = tid; k < n; k += T) {
      d_row[k] = k == i ? 0 : INT_MAX;
      initLocal(k, i);
    }
    __syncthreads();
    if (tid == 0) {
      Q_row[0] = i;
      s_len[0] = 1;
      s_len[1] = s_len[2] = s_len[3] = 0;
      initStack(i);
    }
    __syncthreads();
    if (R[i + 1] == R[i]) continue;  // degree-0 source: its row stays as initialised
    while (true) {
      if (tid == 0) s_next = nwaves;
      __syncthreads();
      const int qlen = s_len[0];
      for (int k = wid; k < qlen;) {
        const int v = Q_row[k];
        const int re = R[v + 1];
        for (int r = R[v] + lane; r < re; r += kWave) visitVertex(d_row, v, C[r], Q2_row, &s_len[1], Q2s_row, &s_len[3]);
        int nk = 0;
        if (lane == 0) nk = atomicAdd(&s_next, 1);
        k = __builtin_amdgcn_readfirstlane(nk);
      }
      __syncthreads();
      const int qslen = s_len[2];
      for (int z = tid; z < qslen; z += T) {
        const int v = Qs_row[z];
        for (int r = R[v]; r < R[v + 1]; r++) visitVertex(d_row, v, C[r], Q2_row, &s_len[1], Q2s_row, &s_len[3]);
      }
      __syncthreads();
      const int q2 = s_len[1], q2s = s_len[3];
      if (q2 == 0 && q2s == 0) {
        for (int k = tid; k < n; k += T) getMax(d_row[k]);
        break;
      }
      for (int kk = tid; kk < q2; kk += T) insertStack(Q2_row, kk);
      for (int kk = tid; kk < q2s; kk += T) insertStack(Q2s_row, kk);
      __syncthreads();
      int *t = Q_row;
      Q_row = Q2_row;
      Q2_row = t;
      t = Qs_row;
      Qs_row = Q2s_row;
      Q2s_row = t;
      if (tid == 0) {
        updateEndpoints(q2);
        s_len[0] = q2;
        s_len[1] = 0;
        s_len[2] = q2s;
        s_len[3] = 0;
      }
      __syncthreads();
    }
    __syncthreads();
    dependencyAccum(d_row, i, tid, lane);
  }
}

}  // namespace gu

#endif
