// gu_engine.hip -- kernels and host orchestration of the batched multi-source
// search (BFS distances, path counts, Brandes dependencies, diameter) and the
// device-wide load-balanced search, behind the C ABI in graph_utils_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/graph_utils_amd.h"
#include "gu_kernels.hpp"
#include "gu_persource.hpp"
#include "gu_internal.hpp"

namespace gu {

// ---- small utility kernels ------------------------------------------------

__global__ void k_fill_u64(u64 *p, size_t count, u64 v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += st) p[i] = v;
}
__global__ void k_fill_i32(int *p, size_t count, int v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += st) p[i] = v;
}

// F[k] = row of entry k (device_graph's COO array when the caller has none).
__global__ void k_build_F(const int *R, int n, int *F) {
  int v = blockIdx.x * (blockDim.x / kWave) + wave_id();
  int stride = gridDim.x * (blockDim.x / kWave);
  for (; v < n; v += stride)
    for (int k = R[v] + lane_id(); k < R[v + 1]; k += kWave) F[k] = v;
}

// Graph validation: R non-decreasing from 0 to m, C in [0, n).
__global__ void k_validate(const int *R, const int *C, int n, long long m, int *bad) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t v = i; v < (size_t)n; v += st)
    if (R[v + 1] < R[v]) atomicOr(bad, 1);
  for (size_t k = i; k < (size_t)m; k += st)
    if (C[k] < 0 || C[k] >= n) atomicOr(bad, 2);
  if (i == 0 && (R[0] != 0 || (long long)R[n] != m)) atomicOr(bad, 4);
}

// ---- batch seeding (shuffle_based.cuh:104-153 per-source initialisation) --
// Clears a batch's state in one launch: lane masks and the three frontier
// slots to 0, claim tags to -1, hub found-masks to 0 and, with path counts,
// the hubs' rows (hubs take their path counts by atomic adds).
template <int NW>
__global__ __launch_bounds__(256) void k_begin(Eng<NW> e, bool sigma) {
  const size_t n = (size_t)e.n;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t i = i0; i < n * NW; i += st) e.visit[i] = 0;
  for (size_t i = i0; i < 3 * n * NW; i += st) e.fr[i] = 0;
  for (size_t i = i0; i < n; i += st) e.tag[i] = -1;
  for (size_t i = i0; i < (size_t)e.nhubs * NW; i += st) e.hub_found[i] = 0;
  if (sigma)
    for (size_t i = i0; i < (size_t)e.nhubs * (kWL * NW); i += st)
      e.svm[(size_t)e.hubs[i / (kWL * NW)] * (kWL * NW) + i % (kWL * NW)] = 0;
}

template <int NW>
__global__ __launch_bounds__(256) void k_seed(Eng<NW> e) {
  __shared__ u64 s_deg[4];
  const int tid = threadIdx.x;
  u64 *f0 = frs(e, 0);
  u64 deg = 0;
  if (tid < e.nb) {
    int s = e.srcs[tid];
    int j = tid / kWL;
    u64 bit = 1ull << (tid % kWL);
    atomicOr(&e.visit[(size_t)s * NW + j], bit);
    atomicOr(&f0[(size_t)s * NW + j], bit);
    e.dvm[(size_t)s * (kWL * NW) + tid] = 0;
    e.svm[(size_t)s * (kWL * NW) + tid] = 1;
    deg = (u64)(e.R[s + 1] - e.R[s]);
  }
  // every lane reaches its source: the first reached vertex and TEPS units
  for (int o = 32; o >= 1; o >>= 1) deg += __shfl_xor(deg, o, kWave);
  if (lane_id() == 0) s_deg[wave_id()] = deg;
  __syncthreads();
  if (tid == 0) {
    e.ctrl->teps = s_deg[0] + s_deg[1] + s_deg[2] + s_deg[3];
    e.ctrl->reached = (u64)e.nb;
  }
  if (tid < e.nb) {
    int s = e.srcs[tid];
    if (atomicMax(&e.tag[s], 0) < 0) {  // first lane with this source pushes it
      unsigned pos = atomicAdd(&e.ctrl->qlen[0], 1u);
      e.q[pos] = s;  // slot 0 and stack offset 0 coincide
      atomicAdd(&e.ctrl->fr_edges[0], (u64)(e.R[s + 1] - e.R[s]));
    }
  }
  __syncthreads();
  if (e.stack) {
    unsigned q0 = atomicAdd(&e.ctrl->qlen[0], 0u);
    if (tid == 0) {
      e.ep[0] = 0;
      e.ep[1] = q0;
      e.ctrl->S_len = q0;
    }
  }
}

// ---- one level ------------------------------------------------------------
template <int NW, int VAR, bool SIGMA>
__global__ __launch_bounds__(1024) void k_expand(Eng<NW> e) {
  if (e.ctrl->done) return;
  const int L = e.ctrl->level;
  const int mode = e.ctrl->mode;
  const unsigned nt = gridDim.x * blockDim.x;
  const unsigned gt = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= 1) {  // clear the stale frontier slot (level L-1) for reuse at L+1
    const int *Qs = qptr(e, L - 1);
    const unsigned qs = e.ctrl->qlen[slot_of(L - 1)];
    u64 *fs = frs(e, L - 1);
    for (unsigned i = gt; i < qs; i += nt) {
      int v = Qs[i];
#pragma unroll
      for (int j = 0; j < NW; j++) fs[(size_t)v * NW + j] = 0;
    }
  }
  if (e.stack) {  // lanes of every stack entry of this level's queue
    const int *Q = qptr(e, L);
    const unsigned qn = e.ctrl->qlen[slot_of(L)];
    const u64 *fc = frs(e, L);
    const long long b = e.ep[L];
    for (unsigned i = gt; i < qn; i += nt) {
      int v = Q[i];
#pragma unroll
      for (int j = 0; j < NW; j++) e.smask[((size_t)b + i) * NW + j] = fc[(size_t)v * NW + j];
    }
  }
  Pusher pb{wave_pushbuf<NW>(), 0};
  Acc a;
  if (mode != 0) {
    // bottom-up hub segments read only level-L rows and write only hub
    // partials, so they run in this launch, ahead of the vertex blocks (one
    // launch fewer per level: 3% per pass on R-MAT s20)
    hub_segments<NW, SIGMA>(e, L, mode, wave_group<NW>(), a);
    bu_level<NW, SIGMA>(e, L, pb, wave_group<NW>(), a);
  } else {
    if (VAR == V_SHUFFLE) td_shuffle<NW>(e, L, pb);
    else if (VAR == V_WARP) td_warp<NW>(e, L, pb);
    else if (VAR == V_CTA_WARP) td_cta_warp<NW>(e, L, pb);
    else if (VAR == V_SCAN) td_scan<NW>(e, L, pb);
    else if (VAR == V_EDGE_PARALLEL) td_edge_parallel<NW>(e, L, pb);
    else td_linear<NW>(e, L);
  }
  pb_flush<NW>(e, L, pb);
  write_acct<NW>(e, 0, a);
}

template <int NW, bool SIGMA>
__global__ __launch_bounds__(1024) void k_finalize(Eng<NW> e) {
  if (e.ctrl->done) return;
  const int L = e.ctrl->level;
  const int mode = e.ctrl->mode;
  Acc a;
  if (mode == 0) {  // bottom-up levels ran their hub segments in k_expand
    td_finalize<NW, SIGMA>(e, L, wave_group<NW>(), a);
    // top-down hub segments only carry path counts (the expansion already
    // pushed the hubs' new lanes); at level 0 the sources push them instead
    if (SIGMA && L == 0)
      hub_push_sources<NW>(e, a);
    else if (SIGMA)
      hub_segments<NW, SIGMA>(e, L, mode, wave_group<NW>(), a);
  }
  // The hub commit touches only hubs' masks, distances and pushes: top-down
  // it reads the lanes the expansion pushed, bottom-up the lanes the
  // segments in k_expand found, so it needs nothing computed above and
  // shares this launch (0.9% per s20 pass over a launch of its own).
  Pusher pb{wave_pushbuf<NW>(), 0};
  hub_commit<NW>(e, L, mode, pb, a);
  pb_flush<NW>(e, L, pb);
  write_acct<NW>(e, 1, a);
}

// One workgroup: sum the level's accounting slots, then terminate or rotate
// the queues and choose the next level's direction.
template <int NW>
__global__ __launch_bounds__(256) void k_control(Eng<NW> e) {
  __shared__ u64 s_red[4][kAcctW];
  Ctrl *c = e.ctrl;
  if (c->done) return;
  u64 t[kAcctW] = {};
  // every slot a thread sums is loaded before any is added (8 per step), so
  // the reduction costs two or three L2 round trips, not one per slot
  constexpr int U = 8;
  for (int kid = 0; kid < 3; kid++) {
    const int nk = e.grid_k[kid];
    for (int b0 = threadIdx.x; b0 < nk; b0 += U * blockDim.x) {
      u64 x[U][kAcctW];
#pragma unroll
      for (int i = 0; i < U; i++) {
        const int b = b0 + i * blockDim.x;
        const u64 *slot = e.acct + ((size_t)kid * kAcctMax + (b < nk ? b : 0)) * kAcctW;
#pragma unroll
        for (int c = 0; c < kAcctW; c++) x[i][c] = b < nk ? slot[c] : 0ull;
      }
#pragma unroll
      for (int i = 0; i < U; i++)
#pragma unroll
        for (int c = 0; c < kAcctW; c++) t[c] = c == 3 ? (t[c] | x[i][c]) : t[c] + x[i][c];
    }
  }
  for (int o = 32; o >= 1; o >>= 1)
    for (int i = 0; i < kAcctW; i++) {
      u64 y = __shfl_xor(t[i], o, kWave);
      t[i] = i == 3 ? (t[i] | y) : t[i] + y;
    }
  if (lane_id() == 0)
    for (int i = 0; i < kAcctW; i++) s_red[wave_id()][i] = t[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < kAcctW; i++) {
    t[i] = 0;
    for (int w = 0; w < (int)blockDim.x / kWave; w++) t[i] = i == 3 ? (t[i] | s_red[w][i]) : t[i] + s_red[w][i];
  }
  const int L = c->level;
  const unsigned qn = c->qlen[slot_of(L + 1)];
  c->unvis_edges -= t[1];
  c->dag += t[2];
  c->teps += t[4];
  c->reached += t[5];
  if (t[3]) c->overflow = 1;
  if (c->mode == 0)
    c->td_levels++;
  else
    c->bu_levels++;
  if (qn == 0) {
    c->done = 1;
    c->last_level = L;
    return;
  }
  if (e.stack) {
    e.ep[L + 2] = e.ep[L + 1] + qn;
    c->S_len = e.ep[L + 2];
  }
  c->fr_edges[slot_of(L + 1)] = t[0];
  c->level = L + 1;
  c->qlen[slot_of(L + 2)] = 0;
  for (int i = 0; i < 4; i++) c->work[i] = 0;
  const u64 mf = t[0];
  const u64 mu = c->unvis_edges;
  int mode = c->mode;
  if (c->direction == GU_DIR_TOPDOWN)
    mode = 0;
  else if (c->direction == GU_DIR_BOTTOMUP)
    mode = 1;
  else if (mode == 0) {
    if (mf * (u64)c->alpha > mu) mode = 1;
  } else {
    if ((u64)qn * (u64)c->beta < (u64)c->n) mode = 0;
  }
  c->mode = mode;
}

// ---- device-wide exclusive scan ------------------------------------------
// Loader: value(i) for i < count().  Tile = blockDim * kScanIPT items.
constexpr int kScanIPT = 8;

struct FrontierDeg {
  const int *R;
  const int *Q;
  __device__ int value(long long i) const {
    int v = Q[i];
    return R[v + 1] - R[v];
  }
};
struct Counts {
  const int *c;
  __device__ int value(long long i) const { return c[i]; }
};

// Phase 1: per-tile sums.
template <class Ld>
__device__ void scan_reduce(Ld ld, long long count, long long *tsum) {
  __shared__ long long s_red[16];
  const int T = blockDim.x;
  const long long TS = (long long)T * kScanIPT;
  for (long long tile = blockIdx.x; tile * TS < count; tile += gridDim.x) {
    long long b = tile * TS + (long long)threadIdx.x * kScanIPT;
    long long s = 0;
    for (int k = 0; k < kScanIPT; k++)
      if (b + k < count) s += ld.value(b + k);
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (lane_id() == 0) s_red[wave_id()] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long t = 0;
      for (int w = 0; w < T / kWave; w++) t += s_red[w];
      tsum[tile] = t;
    }
    __syncthreads();
  }
}

// Phase 2: one workgroup scans the tile sums in place; *total = sum.
__device__ void scan_tiles(long long *tsum, long long ntiles, long long *total) {
  __shared__ long long s_carry;
  __shared__ long long s_w[17];
  const int T = blockDim.x;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (long long b = 0; b < ntiles; b += T) {
    long long i = b + threadIdx.x;
    long long x = i < ntiles ? tsum[i] : 0;
    // wave inclusive scan (64-bit, shuffles)
    long long inc = x;
    for (int o = 1; o < kWave; o <<= 1) {
      long long y = __shfl_up(inc, o, kWave);
      if (lane_id() >= o) inc += y;
    }
    if (lane_id() == kWave - 1) s_w[wave_id()] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long acc = 0;
      for (int w = 0; w < T / kWave; w++) {
        long long t = s_w[w];
        s_w[w] = acc;
        acc += t;
      }
      s_w[16] = acc;
    }
    __syncthreads();
    if (i < ntiles) tsum[i] = s_carry + s_w[wave_id()] + inc - x;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += s_w[16];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

// Phase 3: per-tile exclusive scan plus the tile's offset; with write_total,
// out[count] = total as well.
template <class Ld>
__device__ void scan_apply(Ld ld, long long count, const long long *tsum, int *out,
                           const long long *total, bool write_total) {
  __shared__ int s_scan[17];
  const int T = blockDim.x;
  const long long TS = (long long)T * kScanIPT;
  for (long long tile = blockIdx.x; tile * TS < count; tile += gridDim.x) {
    long long b = tile * TS + (long long)threadIdx.x * kScanIPT;
    int vals[kScanIPT];
    int s = 0;
    for (int k = 0; k < kScanIPT; k++) {
      vals[k] = b + k < count ? ld.value(b + k) : 0;
      s += vals[k];
    }
    int tot;
    int pre = block_exclusive_sum(s, s_scan, &tot);
    long long base = tsum[tile] + pre;
    for (int k = 0; k < kScanIPT; k++) {
      if (b + k < count) out[b + k] = (int)base;
      base += vals[k];
    }
  }
  if (write_total && blockIdx.x == 0 && threadIdx.x == 0) out[count] = (int)*total;
}

// Frontier degree scan for the edge-parallel strategy (only when the level
// runs top-down).
template <int NW>
__global__ __launch_bounds__(1024) void k_fscan_reduce(Eng<NW> e) {
  if (e.ctrl->done || e.ctrl->mode != 0) return;
  const int L = e.ctrl->level;
  FrontierDeg ld{e.R, qptr(e, L)};
  scan_reduce(ld, e.ctrl->qlen[slot_of(L)], e.tsum);
}
template <int NW>
__global__ __launch_bounds__(1024) void k_fscan_tiles(Eng<NW> e) {
  if (e.ctrl->done || e.ctrl->mode != 0) return;
  const int L = e.ctrl->level;
  const long long TS = (long long)blockDim.x * kScanIPT;  // same block size as the others
  long long nt = ((long long)e.ctrl->qlen[slot_of(L)] + TS - 1) / TS;
  scan_tiles(e.tsum, nt, e.tsum + (size_t)e.n + 1);
}
template <int NW>
__global__ __launch_bounds__(1024) void k_fscan_apply(Eng<NW> e) {
  if (e.ctrl->done || e.ctrl->mode != 0) return;
  const int L = e.ctrl->level;
  FrontierDeg ld{e.R, qptr(e, L)};
  scan_apply(ld, e.ctrl->qlen[slot_of(L)], e.tsum, e.off, e.tsum + (size_t)e.n + 1, true);
}

// LBS scan kernels over an explicit counts array.
__global__ __launch_bounds__(1024) void k_cscan_reduce(const int *c, long long S, long long *tsum) {
  scan_reduce(Counts{c}, S, tsum);
}
__global__ __launch_bounds__(1024) void k_cscan_tiles(long long *tsum, long long nt, long long *total) {
  scan_tiles(tsum, nt, total);
}
__global__ __launch_bounds__(1024) void k_cscan_apply(const int *c, long long S, const long long *tsum,
                                                      int *out, const long long *total) {
  if (*total >= INT_MAX) return;  // reported as a capacity error by the host
  scan_apply(Counts{c}, S, tsum, out, total, false);
}

// load_balanced_search.cu:175-191 as a device-wide merge: result[k] =
// upper_bound(scan, k) - 1 for k < W, in wave tiles of kLbsXT items (64
// lanes x 4 items x kLbsXR steps, 16-byte stores when the result is 16-byte
// aligned, no workgroup barrier).  k_lbs_tiles first finds every tile's
// first segment, one binary search per thread, so tile t spans segments
// [tlo[t], tlo[t+1]]; a tile inside one segment is a plain fill, and
// otherwise the wave stages those scan entries in its LDS window.  2^26
// items in 100,000 Zipf segments: 72.0 us for the five launches, against
// 103.9 us with workgroup tiles (two barriers each) whose first and last
// segments one thread searched in the global scan; 10^6 segments: 103
// against 155 us.  (A coarse index of the scan in every workgroup's LDS
// for the tile-start searches was slower: 87 against 81 us.)
// steps per tile beyond kLbsSmall segments: 100,000 Zipf segments took
// 72.0 us with 2, against 76.6 with 1 and 81.3 with 4 (10^6 segments: 103,
// 102 and 124 us)
constexpr int kLbsXR = 2;
constexpr long long kLbsXT = (long long)kWave * 4 * kLbsXR;
constexpr int kLbsWR = 4;  // steps per tile of the one-launch form
constexpr long long kLbsWT = (long long)kWave * 4 * kLbsWR;
constexpr int kLbsWin = 512;  // scan entries a wave stages for one tile

__global__ __launch_bounds__(256) void k_lbs_tiles(const int *scan, long long S, const long long *total,
                                                   long long cap, int *tlo) {
  const long long W = *total;
  if (W > cap || W >= INT_MAX) return;  // reported as a capacity error by the host
  const long long nt = (W + kLbsXT - 1) / kLbsXT;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t <= nt; t += stride)
    tlo[t] = t == nt ? (int)S - 1 : seg_search(scan, 0, (int)S - 1, t * kLbsXT);
}

__global__ __launch_bounds__(256) void k_lbs_expand(const int *scan, long long S, const long long *total,
                                                    long long cap, const int *tlo, int *result) {
  const long long W = *total;
  if (W > cap || W >= INT_MAX) return;  // reported as a capacity error by the host
  __shared__ int s_win[256 / kWave][kLbsWin];
  const int lane = lane_id(), wid = wave_id();
  int *win = s_win[wid];
  const int WPB = blockDim.x / kWave;
  const long long nwave = (long long)gridDim.x * WPB;
  const bool v4 = (reinterpret_cast<uintptr_t>(result) & 15) == 0;
  for (long long tile = (long long)blockIdx.x * WPB + wid; tile * kLbsXT < W; tile += nwave) {
    const long long e0 = tile * kLbsXT;
    const long long e1 = e0 + kLbsXT < W ? e0 + kLbsXT : W;
    const int lo = tlo[tile];
    const int hi = tlo[tile + 1];  // at least the segment of item e1 - 1
    const bool one = lo == hi || (long long)scan[lo + 1] > e1 - 1;
    const int nwin = hi - lo + 1;
    const bool staged = !one && nwin <= kLbsWin;
    if (staged) {
      __builtin_amdgcn_wave_barrier();  // the previous tile's window reads are done
      for (int i = lane; i < nwin; i += kWave) win[i] = scan[lo + i];
      __builtin_amdgcn_wave_barrier();
    }
    auto seg = [&](long long k) -> int {
      if (one) return lo;
      return staged ? lo + seg_search(win, 0, nwin - 1, k) : seg_search(scan, lo, hi, k);
    };
#pragma unroll
    for (int r = 0; r < kLbsXR; r++) {
      const long long b = e0 + ((long long)r * kWave + lane) * 4;
      if (v4 && b + 4 <= e1) {
        *reinterpret_cast<int4 *>(result + b) = make_int4(seg(b), seg(b + 1), seg(b + 2), seg(b + 3));
      } else {
        for (int q = 0; q < 4; q++)
          if (b + q < e1) result[b + q] = seg(b + q);
      }
    }
  }
}

// Load-balanced search for up to kLbsSmall segments in one launch: every
// workgroup scans the counts itself in LDS (a few KB, read from L2), so the
// three scan launches and their dependency disappear; workgroup 0 also
// writes the scan and the total.  Larger inputs take the three-phase scan
// and k_lbs_expand.
// The expansion is a write stream: each wave takes tiles of 64 lanes x 4
// items x kLbsWR steps with no workgroup barrier, finds the tile's first
// and last segment by a wave search, and lanes store 4 items at a time (16 bytes when
// the result is 16-byte aligned).  Against workgroup tiles with two
// barriers and one-item stores, 2^26 items in Zipf segments went from 1090
// to 1220-1237 G items/s.  Measured and slower: 2 or 8 steps per tile
// (887, 1154), 2 or 8 workgroups per CU (848, 1141), non-temporal stores
// (1086), contiguous tile runs per wave (1140), and the last segment
// searched from the first with a one-read check (1196).
// Largest i in [lo, lo + n) with off[i] <= x (off nondecreasing, off[lo] <=
// x), found by the whole wave: 64 evenly spaced samples and a ballot narrow
// the range 64-fold a step, so 8192 entries take three LDS reads per lane.
// Every lane returns the same value.
__device__ __forceinline__ int wave_seg_search(const int *off, int lo, int n, long long x) {
  const int lane = lane_id();
  while (n > kWave) {
    const int step = (n + kWave - 1) / kWave;
    const int idx = lo + lane * step;
    const bool le = idx < lo + n && (long long)off[idx] <= x;
    const int c = __popcll(ballot(le));  // samples at or below x form a prefix
    const int nlo = lo + (c - 1) * step;
    n = (lo + n - nlo) < step ? (lo + n - nlo) : step;
    lo = nlo;
  }
  const bool le = lane < n && (long long)off[lo + lane] <= x;
  return lo + __popcll(ballot(le)) - 1;
}

constexpr int kLbsSmall = 8192;
constexpr int kLbsSmallT = 512;  // threads per workgroup, two per CU
__global__ __launch_bounds__(kLbsSmallT) void k_lbs_small(const int *counts, int S, int *scan, long long cap,
                                                          int *result, long long *total_out) {
  __shared__ int s_off[kLbsSmall];
  __shared__ long long s_wsum[kLbsSmallT / kWave];
  const int tid = threadIdx.x, T = blockDim.x, lane = lane_id(), wid = wave_id();
  const int per = (S + T - 1) / T;
  const int i0 = tid * per < S ? tid * per : S;
  const int i1 = i0 + per < S ? i0 + per : S;
  for (int i = tid; i < S; i += T) s_off[i] = counts[i];  // coalesced
  __syncthreads();
  long long acc = 0;
  for (int i = i0; i < i1; i++) acc += s_off[i];
  long long inc = acc;
  for (int o = 1; o < kWave; o <<= 1) {
    const long long y = __shfl_up(inc, o, kWave);
    if (lane >= o) inc += y;
  }
  if (lane == kWave - 1) s_wsum[wid] = inc;
  __syncthreads();
  long long run = inc - acc, W = 0;
  for (int w = 0; w < T / kWave; w++) {
    if (w < wid) run += s_wsum[w];
    W += s_wsum[w];
  }
  for (int i = i0; i < i1; i++) {
    const int c = s_off[i];
    s_off[i] = (int)run;  // prefixes fit in int whenever W < INT_MAX
    run += c;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int i = tid; i < S; i += T) scan[i] = s_off[i];
    if (tid == 0) *total_out = W;
  }
  if (W > cap || W >= INT_MAX) return;  // reported as a capacity error by the host
  constexpr long long WT = kLbsWT;
  const int WPB = T / kWave;
  const long long nwave = (long long)gridDim.x * WPB;
  const bool v4 = (reinterpret_cast<uintptr_t>(result) & 15) == 0;
  for (long long tile = (long long)blockIdx.x * WPB + wid; tile * WT < W; tile += nwave) {
    const long long e0 = tile * WT;
    const long long e1 = e0 + WT < W ? e0 + WT : W;
    // the tile's first and last segment, by the whole wave (lane 0 alone,
    // with two binary searches, took 50.2 against 48.6 us per pass)
    const int lo = wave_seg_search(s_off, 0, S, e0);
    const int hi = wave_seg_search(s_off, lo, S - lo, e1 - 1);
#pragma unroll
    for (int r = 0; r < kLbsWR; r++) {
      const long long b = e0 + ((long long)r * kWave + lane) * 4;
      if (v4 && b + 4 <= e1) {
        int4 q = make_int4(lo, lo, lo, lo);
        if (lo != hi) {
          q.x = seg_search(s_off, lo, hi, b);
          q.y = seg_search(s_off, q.x, hi, b + 1);
          q.z = seg_search(s_off, q.y, hi, b + 2);
          q.w = seg_search(s_off, q.z, hi, b + 3);
        }
        *reinterpret_cast<int4 *>(result + b) = q;
      } else {
        for (int q = 0; q < 4; q++)
          if (b + q < e1) result[b + q] = lo == hi ? lo : seg_search(s_off, lo, hi, b + q);
      }
    }
  }
}

// ---- outputs ----------------------------------------------------------------
// Transpose one vertex-major per-lane array of the batch ([n][64*NW]) into the
// caller's source-major rows ([nb][n]); lanes that did not reach v get
// `fill`.  A workgroup moves 64 vertices x 64 lanes through LDS; every
// thread keeps 16 loads in flight.
// Output transposes stream: rows are read and written once (non-temporal
// loads and stores measured 0.7% faster per pass than cached ones; issuing
// the next tile's loads before storing the current one gained nothing).
// Loading each value only after its vertex's lane mask, and storing it only
// after loading its row pointer, ran at 3.3 TB/s; with the three loads side
// by side a standalone probe of this tile moves 5.5 TB/s.
constexpr int kTransposeWgPerCu = 8;
template <int NW, class T>
__global__ __launch_bounds__(256) void k_transpose_out(Eng<NW> e, const T *src, T *const *rows, T fill) {
  __shared__ T tile[64][65];
  const int tid = threadIdx.x;
  const int col = tid & 63;
  const int row0 = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index: uniform
  const int words = (e.nb + kWL - 1) / kWL;
  const long long ntiles = (((long long)e.n + 63) / 64) * words;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long long v0 = (t / words) * 64;
    const int j = (int)(t % words);
    // the values, the lane masks of their vertices and the destination rows
    // are loaded side by side; no load waits on another
    T x[16];
    u64 vm[16];
    gp<T> rp[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const long long v = v0 + row0 + 4 * k;
      const bool in = v < e.n;
      x[k] = in ? __builtin_nontemporal_load(&src[(size_t)v * (kWL * NW) + j * kWL + col]) : fill;
      vm[k] = in ? e.visit[(size_t)v * NW + j] : 0ull;
      const int b = j * kWL + row0 + 4 * k;
      rp[k] = glob(b < e.nb ? rows[b] : nullptr);
    }
#pragma unroll
    for (int k = 0; k < 16; k++) tile[row0 + 4 * k][col] = ((vm[k] >> col) & 1ull) ? x[k] : fill;
    __syncthreads();
    const long long v = v0 + col;
    if (v < e.n) {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        gp<T> r = rp[k];
        if (r) __builtin_nontemporal_store(tile[col][row0 + 4 * k], &r[v]);
      }
    }
    __syncthreads();
  }
}

// Eccentricity of every lane: the largest distance among its reached
// vertices (diameter_sampling's per-source atomicMax over finite d,
// shuffle_based.cu:82-111).  Block partial maxima go to ecc with one atomic
// per lane per block; ecc starts at zero.
template <int NW>
__global__ __launch_bounds__(256) void k_lane_ecc(Eng<NW> e, int *ecc) {
  __shared__ int s_m[4][kWL * NW];
  const int tid = threadIdx.x, col = tid & 63, r0 = tid >> 6;
  int mx[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) mx[j] = 0;
  for (long long v = (long long)blockIdx.x * 4 + r0; v < e.n; v += (long long)gridDim.x * 4) {
#pragma unroll
    for (int j = 0; j < NW; j++) {
      if ((e.visit[(size_t)v * NW + j] >> col) & 1ull)
        mx[j] = max(mx[j], e.dvm[(size_t)v * (kWL * NW) + j * kWL + col]);
    }
  }
#pragma unroll
  for (int j = 0; j < NW; j++) s_m[r0][j * kWL + col] = mx[j];
  __syncthreads();
  for (int b = tid; b < e.nb; b += blockDim.x) {
    const int m = max(max(s_m[0][b], s_m[1][b]), max(s_m[2][b], s_m[3][b]));
    if (m > 0) atomicMax(&ecc[b], m);
  }
}

// ---- Brandes backward (shuffle_based.cu:217-354 pull form) ----------------
// Level L, lanes mk of vertex v (lanes for which d(v) = L):
//   delta_b(v) = sum over neighbours w with d_b(w) = L+1 of
//                (float)sigma_b(v) * q_b(w),  q_b(w) = (1 + delta_b(w)) / (float)sigma_b(w)
// summed in FP32 with Neumaier compensation.  q is stored when delta(w) is
// final, so a successor costs one FP32 row instead of its u64 path count and
// FP32 dependency rows (the reference's sigma_v / sigma_w * (1 + delta_w),
// bc.cu:52, reassociated: both are within a few ulp of the FP64 value).  Successor lanes come from a
// dense per-parity mask array tagged with the level that wrote it, so a
// level never needs a clearing pass.
struct Bw {
  float *dl;          // [n][64*NW] dependencies
  float *ql;          // [n][64*NW] (1 + delta) / sigma of each vertex's lanes
                      //   once its delta is final: one row per successor
  int *btag;          // [2][n] level that last wrote each parity's mask
  u64 *blm;           // [2][n][NW] lanes at that level
  unsigned *bwork;    // per-level work counters
  float2 *part;       // [nsegs][64*NW] hub segment partials (sum, compensation)
  const int *hub_seg0;  // [nhubs+1] first segment of each hub
};

__device__ __forceinline__ void neumaier(float &sum, float &comp, float x) {
  float t = sum + x;
  if (fabsf(sum) >= fabsf(x))
    comp += (sum - t) + x;
  else
    comp += (x - t) + sum;
  sum = t;
}

// Accumulate the dependency terms of v's neighbours in [r0, r1) for lanes mk
// (sv = float(sigma_b(v)) per lane).  Up to D successor rows are loaded
// before any is used; the terms are still added in adjacency order.
template <int NW>
__device__ __forceinline__ void bw_accum(const Eng<NW> &e, const Bw &b, int L, const u64 *mk,
                                         const float *sv, int r0, int r1, float *sum, float *comp) {
  const int lane = lane_id();
  const int pr = (L + 1) & 1;
  const int *tagr = b.btag + (size_t)pr * e.n;
  const u64 *lmr = b.blm + (size_t)pr * e.n * NW;
  for (int e0 = r0; e0 < r1; e0 += kWave) {
    int ed = e0 + lane;
    bool valid = ed < r1;
    // branch-free: lanes past the range re-read entry r0, and every lane
    // loads its neighbour's tag and lane masks, selected afterwards (under
    // branches each load waited for the previous one)
    const int w = e.C[valid ? ed : r0];
    const int tg = tagr[w];
    u64 lm[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) lm[j] = lmr[(size_t)w * NW + j];
    const bool ok = valid && tg == L + 1;
    u64 t[NW];
    bool has = false;
#pragma unroll
    for (int j = 0; j < NW; j++) {
      t[j] = ok ? (lm[j] & mk[j]) : 0ull;
      has |= t[j] != 0;
    }
    uint64_t pm = ballot(has);
    while (pm) {
      // successor rows in flight, measured on s22 BC: 16 for 64 lanes (32 no
      // better), 2 for 128 (4: -3%, 8: -8%), 1 for 256 (2: -2%, 4: -5%):
      // wider lanes hold more registers per row, so fewer rows keep occupancy
      constexpr int D = NW == 1 ? 16 : (NW == 2 ? 2 : 1);
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
      float qw[D][NW];
      bool use[D][NW];
#pragma unroll
      for (int u = 0; u < D; u++) {
        if (k[u] >= 0) {
          const int wk = readlane(w, k[u]);
          const size_t rowb = (size_t)wk * (kWL * NW);
#pragma unroll
          for (int j = 0; j < NW; j++) {
            u64 tk = readlane64(t[j], k[u]);
            use[u][j] = (tk >> lane) & 1ull;
            qw[u][j] = use[u][j] ? b.ql[rowb + j * kWL + lane] : 0.0f;
          }
        } else {
#pragma unroll
          for (int j = 0; j < NW; j++) use[u][j] = false;
        }
      }
#pragma unroll
      for (int u = 0; u < D; u++)
#pragma unroll
        for (int j = 0; j < NW; j++)
          if (use[u][j]) neumaier(sum[j], comp[j], sv[j] * qw[u][j]);
    }
  }
}

template <int NW>
__device__ __forceinline__ void bw_lanes(const Eng<NW> &e, int v, int L, u64 *mk) {
  // lanes b with d_b(v) = L, from v's own distance row
  const int lane = lane_id();
  const int *drow = e.dvm + (size_t)v * (kWL * NW);
  // every word is loaded before any is tested (the short-circuit form put a
  // wait after each of the 2*NW dependent loads)
  u64 vis[NW];
  int dl[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    vis[j] = e.visit[(size_t)v * NW + j];
    dl[j] = drow[j * kWL + lane];
  }
#pragma unroll
  for (int j = 0; j < NW; j++) mk[j] = ballot(((vis[j] >> lane) & 1ull) && dl[j] == L);
}

// Hub segments of the backward pass: each wave writes its segment's
// compensated partial sums; the combine kernel adds them in segment order,
// so the result does not depend on scheduling.
template <int NW>
__device__ __forceinline__ void bw_hub_seg(const Eng<NW> &e, const Bw &b, int L, int si) {
  const int lane = lane_id();
  const int4 sg = e.segs[si];
  const int v = sg.x;
  u64 mk[NW];
  bw_lanes<NW>(e, v, L, mk);
  bool any = false;
#pragma unroll
  for (int j = 0; j < NW; j++) any |= mk[j] != 0;
  if (!any) return;
  float sv[NW], sum[NW], comp[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    sum[j] = 0.0f;
    comp[j] = 0.0f;
    const u64 x = e.svm[(size_t)v * (kWL * NW) + j * kWL + lane];  // loaded by every lane
    sv[j] = ((mk[j] >> lane) & 1ull) ? (float)x : 0.0f;
  }
  bw_accum<NW>(e, b, L, mk, sv, sg.y, sg.z, sum, comp);
#pragma unroll
  for (int j = 0; j < NW; j++) b.part[(size_t)si * (kWL * NW) + j * kWL + lane] = make_float2(sum[j], comp[j]);
}

// One backward level.  Waves first take the hub segments (statically: they
// are the long items), then claim 16-entry chunks of the level's stack
// entries; both read only level L+1 rows and write disjoint data (segment
// partials; the rows, masks and tags of non-hub vertices at level L), so
// they share one launch.
template <int NW>
__global__ __launch_bounds__(256) void k_backward(Eng<NW> e, Bw b, int L) {
  const int lane = lane_id();
  const long long beg = e.ep[L], end = e.ep[L + 1];
  const long long cnt_all = end - beg;
  const int pw = L & 1;
  int *tagw = b.btag + (size_t)pw * e.n;
  u64 *lmw = b.blm + (size_t)pw * e.n * NW;
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int si = blockIdx.x * (blockDim.x / kWave) + wave_id(); si < e.nsegs; si += waves)
    bw_hub_seg<NW>(e, b, L, si);
  constexpr int CH = 16;  // stack entries per claim (8 and 32 measured no better)
  while (true) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&b.bwork[L], (unsigned)CH);
    base = bcast(base);
    if ((long long)base >= cnt_all) break;
    const int cnt = cnt_all - base < CH ? (int)(cnt_all - base) : CH;
    for (int i = 0; i < cnt; i++) {
      const long long pos = beg + base + i;
      const int v = e.q[pos];
      const int r0 = e.R[v], r1 = e.R[v + 1];
      if (r1 - r0 > e.hub_deg) continue;  // hubs: segment partials + combine
      u64 mk[NW];
      float sv[NW], sum[NW], comp[NW];
#pragma unroll
      for (int j = 0; j < NW; j++) {
        mk[j] = e.smask[(size_t)pos * NW + j];
        sum[j] = 0.0f;
        comp[j] = 0.0f;
        const u64 x = e.svm[(size_t)v * (kWL * NW) + j * kWL + lane];  // loaded by every lane
        sv[j] = ((mk[j] >> lane) & 1ull) ? (float)x : 0.0f;
      }
      bw_accum<NW>(e, b, L, mk, sv, r0, r1, sum, comp);
      const size_t vrow = (size_t)v * (kWL * NW);
#pragma unroll
      for (int j = 0; j < NW; j++)
        if ((mk[j] >> lane) & 1ull) {
          const float dv = sum[j] + comp[j];
          b.dl[vrow + j * kWL + lane] = dv;
          b.ql[vrow + j * kWL + lane] = (1.0f + dv) / sv[j];
        }
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NW; j++) lmw[(size_t)v * NW + j] = mk[j];
        tagw[v] = L;
      }
    }
  }
}

template <int NW>
__global__ __launch_bounds__(256) void k_bw_hub_combine(Eng<NW> e, Bw b, int L) {
  const int lane = lane_id();
  const int pw = L & 1;
  int *tagw = b.btag + (size_t)pw * e.n;
  u64 *lmw = b.blm + (size_t)pw * e.n * NW;
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int h = blockIdx.x * (blockDim.x / kWave) + wave_id(); h < e.nhubs; h += waves) {
    const int v = e.hubs[h];
    u64 mk[NW];
    bw_lanes<NW>(e, v, L, mk);
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; j++) any |= mk[j] != 0;
    if (!any) continue;
    const size_t vrow = (size_t)v * (kWL * NW);
#pragma unroll
    for (int j = 0; j < NW; j++) {
      float sum = 0.0f, comp = 0.0f;
      for (int si = b.hub_seg0[h]; si < b.hub_seg0[h + 1]; si++) {
        float2 p = b.part[(size_t)si * (kWL * NW) + j * kWL + lane];
        neumaier(sum, comp, p.x);
        comp += p.y;
      }
      if ((mk[j] >> lane) & 1ull) {
        const float dv = sum + comp;
        b.dl[vrow + j * kWL + lane] = dv;
        b.ql[vrow + j * kWL + lane] = (1.0f + dv) / (float)e.svm[vrow + j * kWL + lane];
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < NW; j++) lmw[(size_t)v * NW + j] = mk[j];
      tagw[v] = L;
    }
  }
}

// Dependency rows (source-major, optional) and the FP64 BC accumulation
// bc[v] += sum over lanes of delta_b(v), summed lane by lane in a fixed
// order so the result is deterministic.
template <int NW>
__global__ __launch_bounds__(256) void k_bc_out(Eng<NW> e, const float *dl, float *const *drows,
                                                double *const *bct) {
  __shared__ float tl[64][65];
  const int tid = threadIdx.x, T = blockDim.x;
  const int col = tid % 64;
  for (long long v0 = (long long)blockIdx.x * 64; v0 < e.n; v0 += (long long)gridDim.x * 64) {
    double acc = 0.0;
    for (int j = 0; j < NW; j++) {
      if (j * kWL >= e.nb) break;
      for (int rr = tid / 64; rr < 64; rr += T / 64) {
        long long v = v0 + rr;
        float x = 0.0f;
        if (v < e.n && ((e.visit[(size_t)v * NW + j] >> col) & 1ull)) {
          size_t idx = (size_t)v * (kWL * NW) + j * kWL + col;
          if (e.dvm[idx] != 0) x = dl[idx];  // the source itself has delta 0
        }
        tl[rr][col] = x;
      }
      __syncthreads();
      if (drows) {
        for (int bb = tid / 64; bb < 64; bb += T / 64) {
          int b = j * kWL + bb;
          long long v = v0 + col;
          if (b < e.nb && v < e.n) {
            gp<float> r = glob(drows[b]);
            if (r) r[v] = tl[col][bb];
          }
        }
      }
      if (bct && tid < 64) {
        // lanes in order; a run of lanes with one target is summed, then
        // added to that target once
        for (int c = 0; c < 64; c++) {
          const int b = j * kWL + c;
          if (b >= e.nb) break;
          acc += (double)tl[tid][c];
          gp<double> t = glob(bct[b]);
          if (b + 1 == e.nb || glob(bct[b + 1]) != t) {
            if (t && v0 + tid < e.n) t[v0 + tid] += acc;
            acc = 0.0;
          }
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace gu

// =============================================================================
// Host side
// =============================================================================
using namespace gu;

namespace {

#define GU_HIP(call)                      \
  do {                                    \
    hipError_t err_ = (call);             \
    if (err_ != hipSuccess) return fail(err_); \
  } while (0)

thread_local char g_last_err[256];

// Two events destroyed with the scope, so early error returns leak nothing.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

int fail(hipError_t e) {
  std::snprintf(g_last_err, sizeof(g_last_err), "%s", hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? GU_ERR_NOMEM : GU_ERR_HIP;
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}

struct Workspace {
  int NW = 0;
  int n = 0;
  bool stack = false;
  u64 *visit = nullptr, *fr = nullptr, *svm = nullptr, *smask = nullptr, *blm = nullptr;
  u64 *hub_found = nullptr;
  u64 *acct = nullptr;
  int *q = nullptr, *tag = nullptr, *dvm = nullptr, *off = nullptr, *srcs = nullptr, *btag = nullptr;
  float *dl = nullptr, *ql = nullptr;
  float2 *part = nullptr;
  long long *ep = nullptr, *tsum = nullptr;
  unsigned *bwork = nullptr;
  size_t S_cap = 0;   // entries of q in stack mode
  size_t bytes = 0;
  void release() {
    void *ps[] = {visit, fr, svm, smask, blm, hub_found, acct, q, tag, dvm, off, srcs, btag, dl, ql, part, ep, tsum, bwork};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    *this = Workspace();
  }
};

}  // namespace

// Per-source engine buffers (allocated on first use).
struct PsWorkspace {
  int cap = 0;              // sources per launch
  bool dist = false;        // has a distance scratch of cap rows
  int *queues = nullptr;    // [cap][2n+2] frontier queues / Brandes stack
  int *dscratch = nullptr;  // [cap][n] distances when the caller keeps none
  u64 *sscratch = nullptr;  // [cap][n] path counts of a Brandes launch
  float *xscratch = nullptr;  // [cap][n] dependencies when the caller keeps none
  float *qscratch = nullptr;  // [cap][n] (1 + delta) / sigma of a Brandes launch
  int *depth = nullptr, *ovf = nullptr;
  int *srcs = nullptr;      // [cap] sources of the launch
  u64 *part = nullptr;
  // queue-fed calls (more sources than one launch): per-source tables
  int jcap = 0;             // sources they hold
  int *jsrcs = nullptr;     // [jcap] sources (cluster ids where reordered)
  void **jrows = nullptr;   // [4][jcap] row tables: dist, sigma, ddst, sdst
  u64 *jpart = nullptr;     // [jcap][4]
  int *jdepth = nullptr;    // [jcap]
  int *head = nullptr;      // the queue's counter
  u64 *slog = nullptr;      // [kPsCap][2 * kLdsLog] path-count logs of the level-set kernels
  size_t bytes = 0;
  void release() {
    void *ps[] = {queues, dscratch, sscratch, xscratch, qscratch, depth, ovf, srcs, part,
                  jsrcs, jrows, jpart, jdepth, head, slog};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    *this = PsWorkspace();
  }
};

// Degree-balanced work units of one adjacency (fixed per graph).
struct Units {
  int2 *blocks = nullptr;  // vertex blocks
  int nblocks = 0;
  int4 *segs = nullptr;    // hub adjacency segments (w, e0, e1, hub index)
  int nsegs = 0;
  int *hubs = nullptr;
  int nhubs = 0;
  int *seg0 = nullptr;     // [nhubs+1] first segment of each hub
  bool owned = false;
  void release() {
    if (owned) {
      void *ps[] = {blocks, segs, hubs, seg0};
      for (void *p : ps)
        if (p) (void)hipFree(p);
    }
    *this = Units();
  }
};

struct gu_graph {
  int n = 0;
  long long m = 0;
  int directed = 0;
  int *R = nullptr, *C = nullptr, *F = nullptr;
  int cus = 0;
  int device = 0;
  Workspace ws;
  PsWorkspace pws;
  Ctrl *ctrl = nullptr;       // device
  Ctrl *ctrl_h = nullptr;     // pinned host mirror
  int *srcs_h = nullptr;      // pinned staging for batch sources
  int max_deg = 0;
  double avg_deg = 0;         // over vertices with out-degree > 0
  bool simple = false;        // every adjacency row strictly increasing
  int *RT = nullptr, *CT = nullptr;  // in-edge CSR of a directed graph
  // Cluster-ordered copy of the graph for the per-source engine (bounded-
  // degree graphs): LR/LC in new ids, nof[old] = new id (device and host).
  int *LR = nullptr, *LC = nullptr, *nof = nullptr;
  // Degree <= 4 graphs: the cluster-ordered neighbours as one int4 per
  // vertex (-1 padded), so a per-source visit is one load and a frontier
  // entry needs no row offsets.
  int4 *LE = nullptr;
  // The same neighbours as four int16 deltas from the vertex's own id
  // (kLe8None: none), when every delta fits: half the bytes per visit for
  // the level-set search.
  u64 *LE8 = nullptr;
  int lds_flags = 0;          // kLds* facts of an undirected LE graph
  std::vector<int32_t> nof_h;
  Units pu;                   // work units over in-edges (pulls)
  Units ou;                   // work units over out-edges (Brandes backward)
  int tuned_variant = -1;     // autotuner cache (searches)
  int tuned_direction_alpha = 0;
  // Autotuned launch shape per job family: [0] searches / diameter,
  // [1] Brandes.  Zero fields mean "not tuned".
  struct Tuned {
    int engine = 0, variant = -1, width = 0, wg = 0, alpha = 0, beta = 0;
    int sources = 0;  // launch size it was tuned on
  } tuned[2];
  // Output row tables of one launch (device, with a pinned host mirror),
  // eight tables of 256 entries: [0,256) distance rows, [256,512) path-count
  // rows, [512,768) dependency rows, [768,1024) Brandes accumulation
  // targets, [1024,2048) the caller's rows when the per-source engine works
  // in cluster order.
  void **rowtab = nullptr;
  void **rowtab_h = nullptr;
  int *ecc = nullptr;         // device [256] per-source eccentricity
  int *ecc_h = nullptr;       // pinned mirror
  int *misc_h = nullptr;      // pinned: [0] overflow flag, [64, 320) depths
  u64 *part_h = nullptr;      // pinned [4 * 256] per-source sums and wrap flags
  // Captured chunks of levels of the batched engine, keyed by the launch
  // parameters and the engine arguments (which hold every buffer address);
  // captured on a private stream, replayed on the caller's.
  struct LevelGraph {
    std::vector<char> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<LevelGraph> level_graphs;
  hipStream_t cap_stream = nullptr;
  // Handles made by gu_graph_share read the root's graph arrays and units
  // and own only their workspaces; the root outlives them.
  gu_graph *root = nullptr;
  std::atomic<int> views{0};
};

namespace {

int pick_grid(const gu_graph *g, int per_cu) { return std::max(1, g->cus * per_cu); }

// 64-bit finaliser of splitmix64 (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", 2014): a bijection that mixes every input
// bit into every output bit.
u64 mix64(u64 x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

constexpr int kRowTab = 2048;

void clear_level_graphs(gu_graph *g) {
  for (auto &x : g->level_graphs)
    if (x.exec) (void)hipGraphExecDestroy(x.exec);
  g->level_graphs.clear();
}

// Host copy of the cluster order (a shared handle reads its root's).
const std::vector<int32_t> &nof_host(const gu_graph *g) { return g->root ? g->root->nof_h : g->nof_h; }

// Work-unit shape: a vertex of degree above kHubDeg is a hub whose adjacency
// is pulled in segments of kSeg entries by separate waves; every other
// vertex belongs to a block of at most 64 consecutive vertices holding at
// most kBlockEdges entries (or a single vertex).  Measured on R-MAT s20,
// 64-source BFS + path counts: blocks of 1024 edges and segments of 2048
// ran at 501-510 GTEPS against 435 with 2048 / 1024 (block bounds 512, 768
// and 1280 and segments of 1536-3072 were equal or slower; hubs from 768 or
// 2048 edges no better); sampled BC on s22 gained 2%.
constexpr int kHubDeg = 1024;
constexpr int kSeg = 2048;
constexpr int kBlockEdges = 1024;

int build_units(int n, const int32_t *R, Units *u) {
  std::vector<int2> blocks;
  std::vector<int4> segs;
  std::vector<int> hubs, seg0;
  int b0 = -1, bcount = 0;
  long long bedges = 0;
  auto close = [&](int v) {
    if (b0 >= 0) blocks.push_back(make_int2(b0, v));
    b0 = -1;
    bcount = 0;
    bedges = 0;
  };
  for (int v = 0; v < n; v++) {
    const int d = R[v + 1] - R[v];
    if (d > kHubDeg) {
      close(v);
      const int h = (int)hubs.size();
      hubs.push_back(v);
      seg0.push_back((int)segs.size());
      for (int e0 = R[v]; e0 < R[v + 1]; e0 += kSeg)
        segs.push_back(make_int4(v, e0, std::min(e0 + kSeg, R[v + 1]), h));
      continue;
    }
    if (b0 >= 0 && (bcount == kWave || bedges + d > kBlockEdges)) close(v);
    if (b0 < 0) b0 = v;
    bcount++;
    bedges += d;
  }
  close(n);
  seg0.push_back((int)segs.size());
  u->owned = true;
  u->nblocks = (int)blocks.size();
  u->nsegs = (int)segs.size();
  u->nhubs = (int)hubs.size();
  if (dalloc(&u->blocks, blocks.size()) != hipSuccess || dalloc(&u->segs, segs.size()) != hipSuccess ||
      dalloc(&u->hubs, hubs.size()) != hipSuccess || dalloc(&u->seg0, seg0.size()) != hipSuccess)
    return GU_ERR_NOMEM;
  if ((!blocks.empty() && hipMemcpy(u->blocks, blocks.data(), sizeof(int2) * blocks.size(),
                                    hipMemcpyHostToDevice) != hipSuccess) ||
      (!segs.empty() && hipMemcpy(u->segs, segs.data(), sizeof(int4) * segs.size(),
                                  hipMemcpyHostToDevice) != hipSuccess) ||
      (!hubs.empty() && hipMemcpy(u->hubs, hubs.data(), sizeof(int) * hubs.size(),
                                  hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(u->seg0, seg0.data(), sizeof(int) * seg0.size(), hipMemcpyHostToDevice) != hipSuccess)
    return GU_ERR_HIP;
  return GU_OK;
}

// In-edge CSR (the transpose) of a directed graph, built on the host.
void transpose_csr(int n, long long m, const int32_t *R, const int32_t *C, std::vector<int32_t> &RT,
                   std::vector<int32_t> &CT) {
  RT.assign((size_t)n + 1, 0);
  for (long long k = 0; k < m; k++) RT[(size_t)C[k] + 1]++;
  for (int v = 0; v < n; v++) RT[(size_t)v + 1] += RT[v];
  std::vector<int32_t> pos(RT.begin(), RT.end() - 1);
  CT.assign((size_t)m, 0);
  for (int v = 0; v < n; v++)
    for (int k = R[v]; k < R[v + 1]; k++) CT[(size_t)pos[C[k]]++] = v;
}

// Buffers are sized for the widest batch requested so far (w.NW words per
// vertex); a narrower batch uses the front of each buffer with its own
// [n][64*NW] layout, so mixed widths never reallocate.
int ensure_ws(gu_graph *g, int NW, bool stack, size_t want_S, hipStream_t st) {
  Workspace &w = g->ws;
  const size_t n = (size_t)g->n;
  if (NW > w.NW || (stack && !w.stack)) {
    NW = std::max(NW, w.NW);
    stack = stack || w.stack;
    clear_level_graphs(g);
    w.release();
    w.NW = NW;
    w.n = g->n;
    const size_t WL = (size_t)kWL * NW;
    GU_HIP(dalloc(&w.visit, n * NW));
    GU_HIP(dalloc(&w.fr, 3 * n * NW));
    GU_HIP(dalloc(&w.tag, n));
    GU_HIP(dalloc(&w.dvm, n * WL));
    GU_HIP(dalloc(&w.svm, n * WL));
    GU_HIP(dalloc(&w.off, n + 1));
    GU_HIP(dalloc(&w.tsum, n + 8));
    GU_HIP(dalloc(&w.srcs, 256));
    GU_HIP(dalloc(&w.ep, n + 3));
    GU_HIP(dalloc(&w.hub_found, (size_t)std::max(g->pu.nhubs, 1) * NW));
    GU_HIP(dalloc(&w.acct, (size_t)3 * kAcctMax * kAcctW));
    w.bytes = n * NW * 8 * 4 + n * 4 * 2 + n * WL * 12 + (n + 8) * 8 + 1024 + (n + 3) * 8;
    w.S_cap = 3 * n;
    GU_HIP(dalloc(&w.q, w.S_cap));
    w.bytes += w.S_cap * 4;
    if (stack) {
      w.stack = true;
      GU_HIP(dalloc(&w.smask, w.S_cap * NW));
      GU_HIP(dalloc(&w.dl, n * WL));
      GU_HIP(dalloc(&w.ql, n * WL));
      GU_HIP(dalloc(&w.btag, 2 * n));
      GU_HIP(dalloc(&w.blm, 2 * n * NW));
      GU_HIP(dalloc(&w.bwork, n + 3));
      GU_HIP(dalloc(&w.part, (size_t)std::max(g->ou.nsegs, 1) * WL));
      w.bytes += w.S_cap * NW * 8 + n * WL * 8 + 2 * n * 4 + 2 * n * NW * 8 + (n + 3) * 4;
    }
  }
  if (stack && want_S > w.S_cap) {  // grow the stack, keeping its contents
    size_t cap = std::max(want_S, w.S_cap * 2);
    int *q2;
    u64 *m2;
    GU_HIP(dalloc(&q2, cap));
    if (hipError_t e = dalloc(&m2, cap * w.NW)) {
      (void)hipFree(q2);
      return fail(e);
    }
    hipError_t e = hipMemcpyAsync(q2, w.q, w.S_cap * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(m2, w.smask, w.S_cap * w.NW * 8, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {  // the old stack stays in place
      (void)hipFree(q2);
      (void)hipFree(m2);
      return fail(e);
    }
    (void)hipFree(w.q);
    (void)hipFree(w.smask);
    clear_level_graphs(g);
    w.bytes += (cap - w.S_cap) * (4 + w.NW * 8);
    w.q = q2;
    w.smask = m2;
    w.S_cap = cap;
  }
  return GU_OK;
}

enum Job { JOB_BFS = 0, JOB_BC = 1, JOB_DIAM = 2 };

// Outputs of one batch: device row tables (entry b = row of the batch's
// b-th source, null = not wanted), per-lane Brandes targets and the
// per-lane eccentricities.
struct BatchOut {
  int *const *dist = nullptr;
  u64 *const *sigma = nullptr;
  float *const *delta = nullptr;
  double *const *bc = nullptr;
  int *ecc = nullptr;
};

struct RunInfo {
  int last_level = 0;
  int td = 0, bu = 0;
  int overflow = 0;
  u64 teps = 0, reached = 0, dag = 0;
};

template <int NW>
Eng<NW> make_eng(gu_graph *g, int nb, bool stack, bool sigma) {
  Workspace &w = g->ws;
  Eng<NW> e{};
  e.R = g->R;
  e.C = g->C;
  e.n = g->n;
  e.simple = g->simple ? 1 : 0;
  e.nb = nb;
  e.visit = w.visit;
  e.fr = w.fr;
  e.q = w.q;
  e.smask = w.smask;
  e.ep = w.ep;
  e.tag = w.tag;
  e.dvm = w.dvm;
  e.svm = w.svm;
  e.ctrl = g->ctrl;
  e.off = w.off;
  e.tsum = w.tsum;
  e.srcs = w.srcs;
  e.RP = g->RT ? g->RT : g->R;
  e.CP = g->CT ? g->CT : g->C;
  e.blocks = g->pu.blocks;
  e.nblocks = g->pu.nblocks;
  e.segs = g->pu.segs;
  e.nsegs = g->pu.nsegs;
  e.hubs = g->pu.hubs;
  e.nhubs = g->pu.nhubs;
  e.hub_deg = kHubDeg;
  e.hub_found = w.hub_found;
  e.acct = w.acct;
  for (int j = 0; j < NW; j++) {
    int lanes = std::min(std::max(nb - j * kWL, 0), kWL);
    e.all[j] = lanes == 64 ? ~0ull : ((1ull << lanes) - 1);
  }
  e.stack = stack ? 1 : 0;
  e.sigma = sigma ? 1 : 0;
  return e;
}

template <int NW, bool SIGMA>
void launch_level(gu_graph *g, const Eng<NW> &e, int variant, int block, int grid, hipStream_t st) {
  if (variant == V_EDGE_PARALLEL) {
    k_fscan_reduce<NW><<<grid, 256, 0, st>>>(e);
    k_fscan_tiles<NW><<<1, 256, 0, st>>>(e);
    k_fscan_apply<NW><<<grid, 256, 0, st>>>(e);
  }
  const size_t lx = level_lds_bytes<NW>(block / kWave), l256 = level_lds_bytes<NW>(4);
  switch (variant) {
    case V_LINEAR: k_expand<NW, V_LINEAR, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
    case V_EDGE_PARALLEL: k_expand<NW, V_EDGE_PARALLEL, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
    case V_WARP: k_expand<NW, V_WARP, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
    case V_SCAN: k_expand<NW, V_SCAN, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
    case V_CTA_WARP: k_expand<NW, V_CTA_WARP, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
    default: k_expand<NW, V_SHUFFLE, SIGMA><<<e.grid_k[0], block, lx, st>>>(e); break;
  }
  k_finalize<NW, SIGMA><<<e.grid_k[1], 256, l256, st>>>(e);
  k_control<NW><<<1, 256, 0, st>>>(e);
}

// The executable graph of K consecutive levels for these launch parameters,
// captured on first use.  Level kernels read the level, direction and
// counters from the control block, so one capture serves every chunk of K
// levels of every batch with the same arguments.
template <int NW>
int level_graph(gu_graph *g, const Eng<NW> &e, int variant, int block, int grid, int K, bool sigma,
                hipGraphExec_t *out) {
  std::vector<char> key(sizeof(Eng<NW>) + 6 * sizeof(int));
  int hdr[6] = {NW, variant, block, grid, K, sigma ? 1 : 0};
  std::memcpy(key.data(), hdr, sizeof(hdr));
  std::memcpy(key.data() + sizeof(hdr), &e, sizeof(Eng<NW>));
  for (auto &x : g->level_graphs)
    if (x.key == key) {
      *out = x.exec;
      return GU_OK;
    }
  if (!g->cap_stream) GU_HIP(hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking));
  hipStream_t cs = g->cap_stream;
  GU_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < K; i++) {
    if (sigma)
      launch_level<NW, true>(g, e, variant, block, grid, cs);
    else
      launch_level<NW, false>(g, e, variant, block, grid, cs);
  }
  hipGraph_t gr = nullptr;
  hipError_t err = hipStreamEndCapture(cs, &gr);
  if (err != hipSuccess) return fail(err);
  hipGraphExec_t ex = nullptr;
  err = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (err != hipSuccess) return fail(err);
  if (g->level_graphs.size() >= 24) {  // bounded cache: drop the oldest
    (void)hipGraphExecDestroy(g->level_graphs.front().exec);
    g->level_graphs.erase(g->level_graphs.begin());
  }
  g->level_graphs.push_back(gu_graph::LevelGraph{std::move(key), ex});
  *out = ex;
  return GU_OK;
}

int default_alpha(const gu_config &c) { return c.alpha > 0 ? c.alpha : 16; }
int default_beta(const gu_config &c) { return c.beta > 0 ? c.beta : 64; }

// One batch of nb <= 64*NW sources through the level loop.
template <int NW>
int run_batch(gu_graph *g, const int *srcs, int nb, Job job, bool sigma, const gu_config &cfg,
              int variant, const BatchOut &out, RunInfo *info, hipStream_t st) {
  const bool stack = job == JOB_BC;
  GuTraceRange tr_pass(stack ? "gu:batch brandes" : "gu:batch search");
  int rc = ensure_ws(g, NW, stack, 0, st);
  if (rc) return rc;
  Workspace &w = g->ws;
  const size_t n = (size_t)g->n;
  const int fgrid = pick_grid(g, 4);
  std::memcpy(g->srcs_h, srcs, sizeof(int) * nb);
  GU_HIP(hipMemcpyAsync(w.srcs, g->srcs_h, sizeof(int) * nb, hipMemcpyHostToDevice, st));
  Ctrl c0{};
  c0.direction = cfg.direction;
  c0.alpha = default_alpha(cfg);
  c0.beta = default_beta(cfg);
  c0.mode = cfg.direction == GU_DIR_BOTTOMUP ? 1 : 0;
  c0.unvis_edges = (u64)g->m;
  c0.n = g->n;
  *g->ctrl_h = c0;
  GU_HIP(hipMemcpyAsync(g->ctrl, g->ctrl_h, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  Eng<NW> e = make_eng<NW>(g, nb, stack, sigma);
  k_begin<NW><<<fgrid, 256, 0, st>>>(e, sigma);
  k_seed<NW><<<1, 256, 0, st>>>(e);
  const int block = cfg.wg_size > 0 ? cfg.wg_size : 256;
  // 64 waves per compute unit, twice what is resident at once: waves take
  // their blocks statically, so the second round of workgroups soaks up the
  // first round's imbalance (s20 BFS: 542 GTEPS against 510 with 32 waves,
  // 455 with 16, 521 with 128)
  const int grid = std::min(kAcctMax, pick_grid(g, std::max(1, 64 * kWave / block)));
  auto set_grids = [&](Eng<NW> &x) {
    x.grid_k[0] = grid;
    x.grid_k[1] = std::min(kAcctMax, pick_grid(g, 8));
    x.grid_k[2] = 0;  // accounting slot of a third level kernel (none)
  };
  set_grids(e);
  // first chunk of 3 levels: on R-MAT the poll after level 1 came too early
  // (s20 BFS and the serving stream about 1% ahead of 2; 4 no better)
  int K = cfg.sync_every > 0 ? cfg.sync_every : 3;
  // stack mode reserves room for every level of a chunk, so its chunks stay short
  const int Kmax = cfg.sync_every > 0 ? cfg.sync_every : (stack ? 8 : 64);
  int levels_launched = 0;
  unsigned peak = 0;
  while (true) {
    if (stack) {
      // every level adds at most n stack entries
      rc = ensure_ws(g, NW, true, (size_t)g->ctrl_h->S_len + (size_t)(K + 2) * n, st);
      if (rc) return rc;
      e = make_eng<NW>(g, nb, stack, sigma);
      set_grids(e);
    }
    GuTraceRange tr_chunk("gu:level chunk");
    if (cfg.level_graphs > 0) {
      hipGraphExec_t ex = nullptr;
      rc = level_graph<NW>(g, e, variant, block, grid, K, sigma, &ex);
      if (rc) return rc;
      GU_HIP(hipGraphLaunch(ex, st));
    } else {
      for (int i = 0; i < K; i++) {
        if (sigma)
          launch_level<NW, true>(g, e, variant, block, grid, st);
        else
          launch_level<NW, false>(g, e, variant, block, grid, st);
      }
    }
    levels_launched += K;
    GU_HIP(hipGetLastError());
    GU_HIP(hipMemcpyAsync(g->ctrl_h, g->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    GU_HIP(hipStreamSynchronize(st));
    if (g->ctrl_h->done) break;
    if (levels_launched > g->n + 2) return GU_ERR_INVALID;  // cannot happen for a valid graph
    // Past the frontier's peak the search is in its tail: short chunks, so
    // few levels are launched after the last non-empty one.
    const unsigned cur = g->ctrl_h->qlen[g->ctrl_h->level % 3];
    peak = std::max(peak, cur);
    K = (cfg.sync_every <= 0 && (u64)cur * 8 < (u64)peak) ? 2 : std::min(K * 2, Kmax);
  }
  const int last = g->ctrl_h->last_level;
  if (job == JOB_BFS && (out.dist || out.sigma))
  {
    GuTraceRange tr("gu:rows out");
    const long long tiles = ((long long)g->n + 63) / 64 * ((nb + kWL - 1) / kWL);
    const int og = (int)std::max(1LL, std::min<long long>(tiles, pick_grid(g, kTransposeWgPerCu)));
    if (out.dist) k_transpose_out<NW, int><<<og, 256, 0, st>>>(e, w.dvm, out.dist, INT_MAX);
    if (out.sigma) k_transpose_out<NW, u64><<<og, 256, 0, st>>>(e, w.svm, out.sigma, 0ull);
  }
  if (out.ecc) {
    k_fill_i32<<<1, 256, 0, st>>>(out.ecc, (size_t)nb, 0);
    k_lane_ecc<NW><<<pick_grid(g, 4), 256, 0, st>>>(e, out.ecc);
  }
  if (job == JOB_BC) {
    k_fill_i32<<<fgrid, 256, 0, st>>>(w.btag, 2 * n, -1);
    k_fill_i32<<<1, 256, 0, st>>>((int *)w.bwork, (size_t)last + 2, 0);
    // the backward walks out-edges: hubs and segments of the out-edge units
    Eng<NW> eb = e;
    eb.segs = g->ou.segs;
    eb.nsegs = g->ou.nsegs;
    eb.hubs = g->ou.hubs;
    eb.nhubs = g->ou.nhubs;
    Bw b{w.dl, w.ql, w.btag, w.blm, w.bwork, w.part, g->ou.seg0};
    const int cgrid = std::max(1, std::min(pick_grid(g, 8), (g->ou.nhubs + 3) / 4));
    GuTraceRange tr("gu:brandes backward");
    for (int L = last; L >= 1; L--) {
      // 16 workgroups per compute unit: s22 BC 305 GTEPS against 292 with 8
      k_backward<NW><<<pick_grid(g, 16), 256, 0, st>>>(eb, b, L);
      if (g->ou.nhubs) k_bw_hub_combine<NW><<<cgrid, 256, 0, st>>>(eb, b, L);
    }
    k_bc_out<NW><<<pick_grid(g, 8), 256, 0, st>>>(e, w.dl, out.delta, out.bc);
  }
  GU_HIP(hipGetLastError());
  if (out.ecc) GU_HIP(hipMemcpyAsync(g->ecc_h, out.ecc, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
  GU_HIP(hipMemcpyAsync(g->ctrl_h, g->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  GU_HIP(hipStreamSynchronize(st));
  info->last_level = last;
  info->td = g->ctrl_h->td_levels;
  info->bu = g->ctrl_h->bu_levels;
  info->overflow = g->ctrl_h->overflow;
  info->teps = g->ctrl_h->teps;
  info->reached = g->ctrl_h->reached;
  info->dag = g->ctrl_h->dag;
  return GU_OK;
}

constexpr int kPsCap = 256;  // sources per per-source launch

int ensure_pws(gu_graph *g, bool need_dist) {
  PsWorkspace &p = g->pws;
  const size_t n = (size_t)g->n;
  if (p.cap == 0) {
    GU_HIP(dalloc(&p.queues, (size_t)kPsCap * (2 * n + 2)));
    GU_HIP(dalloc(&p.depth, kPsCap));
    GU_HIP(dalloc(&p.ovf, 1));
    GU_HIP(dalloc(&p.part, 4 * kPsCap));
    GU_HIP(dalloc(&p.srcs, kPsCap));
    p.cap = kPsCap;
    p.bytes = (size_t)kPsCap * (2 * n + 2) * 4 + kPsCap * 32 + 4;
  }
  if (need_dist && !p.dist) {
    GU_HIP(dalloc(&p.dscratch, (size_t)kPsCap * n));
    p.dist = true;
    p.bytes += (size_t)kPsCap * n * 4;
  }
  return GU_OK;
}

// Host view of a job's outputs: arrays of nsrc device row pointers (null
// array or entry = not wanted), per-source Brandes targets and per-source
// eccentricities (host).
struct HostRows {
  int *const *dist = nullptr;
  u64 *const *sigma = nullptr;
  float *const *delta = nullptr;
  double *const *bc = nullptr;
  int32_t *ecc = nullptr;
  // per source (host): 1 if a path count wrapped modulo 2^64; exact per
  // source on the per-source engine, per launch on the batched engine
  int32_t *wrap = nullptr;
};

int run_persource(gu_graph *g, const int *sources, int nsrc, const HostRows &hr, bool sigma,
                  gu_stats &s, int &dmax, hipStream_t st, int queue = 0, bool allow_lds = true);

// The level-set search for what the call wants: path counts (and distance
// rows), distance rows only, or eccentricities only; e8: the graph has the
// int16-delta neighbour table.
void launch_ps_lds(const PsLdsArgs &c, bool sigma, bool rows, bool e8, int grid, hipStream_t st) {
  if (sigma && e8)
    k_ps_lds<true, true, true><<<grid, 1024, 0, st>>>(c);
  else if (sigma)
    k_ps_lds<true, true, false><<<grid, 1024, 0, st>>>(c);
  else if (rows && e8)
    k_ps_lds<false, true, true><<<grid, 1024, 0, st>>>(c);
  else if (rows)
    k_ps_lds<false, true, false><<<grid, 1024, 0, st>>>(c);
  else if (e8)
    k_ps_lds<false, false, true><<<grid, 1024, 0, st>>>(c);
  else
    k_ps_lds<false, false, false><<<grid, 1024, 0, st>>>(c);
}

// Path-count scratch rows, kPsCap of them.  On a graph with the degree-4
// neighbour table a row holds 2n words: the level-set search with path
// counts keeps a 16-byte record per vertex there (ps_rec); the other
// kernels use the first n words of each row.
int ensure_sscratch(gu_graph *g) {
  PsWorkspace &p = g->pws;
  if (p.sscratch) return GU_OK;
  const size_t n = (size_t)g->n, words = (g->LE ? 2 : 1) * n;
  GU_HIP(dalloc(&p.sscratch, (size_t)kPsCap * words));
  p.bytes += (size_t)kPsCap * words * 8;
  return GU_OK;
}

// The level-set kernels' path-count logs, one pair of levels per workgroup
// (32 MB).
int ensure_slog(gu_graph *g) {
  PsWorkspace &p = g->pws;
  if (p.slog) return GU_OK;
  GU_HIP(dalloc(&p.slog, (size_t)kPsCap * 2 * kLdsLog));
  p.bytes += (size_t)kPsCap * 2 * kLdsLog * 8;
  return GU_OK;
}

// Whether a search of g can run on the LDS level sets: an undirected graph
// with the degree-4 neighbour table, whose ids fit the keyed sets when the
// search counts paths.
bool lds_ok(const gu_graph *g, const int4 *LE, bool sigma) {
  return LE && !g->directed && (!sigma || g->n <= kLdsMaxN);
}

int ensure_jobs(gu_graph *g, int nsrc) {
  PsWorkspace &p = g->pws;
  if (p.jcap >= nsrc) return GU_OK;
  void *ps[] = {p.jsrcs, p.jrows, p.jpart, p.jdepth};
  for (void *q : ps)
    if (q) (void)hipFree(q);
  p.bytes -= (size_t)p.jcap * (4 + 32 + 32 + 4);
  p.jsrcs = nullptr, p.jrows = nullptr, p.jpart = nullptr, p.jdepth = nullptr, p.jcap = 0;
  const int cap = std::max(nsrc, 2 * kPsCap);
  GU_HIP(dalloc(&p.jsrcs, (size_t)cap));
  GU_HIP(dalloc(&p.jrows, 4 * (size_t)cap));
  GU_HIP(dalloc(&p.jpart, 4 * (size_t)cap));
  GU_HIP(dalloc(&p.jdepth, (size_t)cap));
  if (!p.head) GU_HIP(dalloc(&p.head, 1));
  p.jcap = cap;
  p.bytes += (size_t)cap * (4 + 32 + 32 + 4);
  return GU_OK;
}

// Per-source searches of a whole call as one grid: at most kPsCap
// workgroups (one per compute unit) take the call's sources from a device
// queue, so a workgroup whose source finishes early starts the next one
// instead of waiting for the launch's deepest search.  On a
// cluster-ordered graph each workgroup searches in its own scratch rows
// and writes its source's rows back to the caller's order when the search
// ends.  Sources whose LDS level sets overflow are searched again on the
// distance-reading kernel.
int run_persource_queue(gu_graph *g, const int *sources, int nsrc, const HostRows &hr, bool sigma,
                        gu_stats &s, int &dmax, hipStream_t st) {
  const bool rl = g->LR != nullptr;
  int rc = ensure_pws(g, true);
  if (rc) return rc;
  if ((rc = ensure_jobs(g, nsrc))) return rc;
  PsWorkspace &p = g->pws;
  const size_t n = (size_t)g->n;
  if (rl && sigma && (rc = ensure_sscratch(g))) return rc;
  const int32_t *nof = rl ? nof_host(g).data() : nullptr;
  std::vector<int> srcs((size_t)nsrc);
  std::vector<void *> rows(4 * (size_t)nsrc, nullptr);
  for (int j = 0; j < nsrc; j++) {
    srcs[j] = rl ? nof[sources[j]] : sources[j];
    int *dr = hr.dist ? hr.dist[j] : nullptr;
    u64 *sr = sigma && hr.sigma ? hr.sigma[j] : nullptr;
    if (sigma && !sr) return GU_ERR_INVALID;  // path counts need every row
    if (rl) {  // searched in the slot's scratch rows, written back by the workgroup
      rows[2 * (size_t)nsrc + j] = dr;
      rows[3 * (size_t)nsrc + j] = sr;
    } else {  // searched in the caller's rows (the slot's scratch when no distances)
      rows[j] = dr;
      rows[(size_t)nsrc + j] = sr;
    }
  }
  GU_HIP(hipMemcpyAsync(p.jsrcs, srcs.data(), sizeof(int) * nsrc, hipMemcpyHostToDevice, st));
  GU_HIP(hipMemcpyAsync(p.jrows, rows.data(), sizeof(void *) * 4 * nsrc, hipMemcpyHostToDevice, st));
  k_fill_i32<<<1, 64, 0, st>>>(p.ovf, 1, 0);
  k_fill_i32<<<1, 64, 0, st>>>(p.head, 1, 0);
  PsQueue q{nsrc, p.head, p.dscratch, p.sscratch, rl ? g->nof : nullptr,
            rl ? (int *const *)(p.jrows + 2 * (size_t)nsrc) : nullptr,
            rl && sigma ? (u64 *const *)(p.jrows + 3 * (size_t)nsrc) : nullptr};
  void *const *dtab = (void *const *)p.jrows;
  u64 *const *stab = sigma ? (u64 *const *)(p.jrows + (size_t)nsrc) : nullptr;
  const int grid = std::min(nsrc, kPsCap);
  const int4 *LE = rl ? g->LE : nullptr;
  const bool lds = lds_ok(g, LE, sigma);
  if (lds && sigma && (rc = ensure_slog(g))) return rc;
  {
    GuTraceRange tr("gu:per-source queue");
    if (lds) {
      PsLdsArgs c{LE, g->n, p.jsrcs, (int *const *)dtab, stab, p.queues, p.jdepth, p.jpart, p.ovf, g->lds_flags,
                  g->LE8};
      c.q = q;
      c.slog = p.slog;
      launch_ps_lds(c, sigma, hr.dist != nullptr || sigma, g->LE8 != nullptr, grid, st);
    } else {
      PsArgs a{rl ? g->LR : g->R, rl ? g->LC : g->C, LE, g->n, p.jsrcs, (int *const *)dtab, stab, p.queues,
               p.jdepth, p.jpart, p.ovf};
      a.q = q;
      if (LE) {
        if (sigma && !g->directed)
          k_persource<true, true, true><<<grid, 1024, 0, st>>>(a);
        else if (sigma)
          k_persource<true, false, true><<<grid, 1024, 0, st>>>(a);
        else
          k_persource<false, false, true><<<grid, 1024, 0, st>>>(a);
      } else if (sigma && !g->directed) {
        k_persource<true, true><<<grid, 1024, 0, st>>>(a);
      } else if (sigma) {
        k_persource<true, false><<<grid, 1024, 0, st>>>(a);
      } else {
        k_persource<false, false><<<grid, 1024, 0, st>>>(a);
      }
    }
    GU_HIP(hipGetLastError());
  }
  std::vector<u64> part(4 * (size_t)nsrc);
  std::vector<int> depth((size_t)nsrc);
  GU_HIP(hipMemcpyAsync(part.data(), p.jpart, sizeof(u64) * 4 * nsrc, hipMemcpyDeviceToHost, st));
  GU_HIP(hipMemcpyAsync(depth.data(), p.jdepth, sizeof(int) * nsrc, hipMemcpyDeviceToHost, st));
  GU_HIP(hipMemcpyAsync(g->misc_h, p.ovf, sizeof(int), hipMemcpyDeviceToHost, st));
  GU_HIP(hipStreamSynchronize(st));
  s.sigma_overflow |= g->misc_h[0];
  // sources whose level sets overflowed: searched again, rows included
  std::vector<int> again;
  for (int j = 0; j < nsrc; j++) {
    if (lds && (part[4 * (size_t)j + 3] & 2)) {
      again.push_back(j);
      continue;
    }
    s.teps_units += (int64_t)part[4 * (size_t)j];
    s.reached += (int64_t)part[4 * (size_t)j + 1];
    s.dag_edges += (int64_t)part[4 * (size_t)j + 2];
    dmax = std::max(dmax, depth[j]);
    if (hr.ecc) hr.ecc[j] = depth[j];
    if (hr.wrap) hr.wrap[j] = sigma && (part[4 * (size_t)j + 3] & 1) != 0;
  }
  s.batches++;
  if (again.empty()) return GU_OK;
  std::vector<int> asrc(again.size());
  std::vector<int32_t *> ad(again.size());
  std::vector<u64 *> as(again.size());
  std::vector<int32_t> ecc2(again.size()), wrap2(again.size());
  for (size_t i = 0; i < again.size(); i++) {
    asrc[i] = sources[again[i]];
    ad[i] = hr.dist ? hr.dist[again[i]] : nullptr;
    as[i] = hr.sigma ? hr.sigma[again[i]] : nullptr;
  }
  HostRows ar;
  ar.dist = hr.dist ? ad.data() : nullptr;
  ar.sigma = hr.sigma ? as.data() : nullptr;
  ar.ecc = hr.ecc ? ecc2.data() : nullptr;
  ar.wrap = hr.wrap ? wrap2.data() : nullptr;
  rc = run_persource(g, asrc.data(), (int)again.size(), ar, sigma, s, dmax, st, 2, false);
  if (rc) return rc;
  for (size_t i = 0; i < again.size(); i++) {
    if (hr.ecc) hr.ecc[again[i]] = ecc2[i];
    if (hr.wrap) hr.wrap[again[i]] = wrap2[i];
  }
  return GU_OK;
}

// Per-source engine over nsrc sources, kPsCap workgroups per launch.  On a
// cluster-ordered graph the traversal runs in new ids on scratch rows and
// k_ps_unpermute writes the caller's rows.  queue: 0 = the queue-fed grid
// when the sources exceed one launch, 1 = always, 2 = never; allow_lds =
// false keeps the search on the distance-reading kernels.
int run_persource(gu_graph *g, const int *sources, int nsrc, const HostRows &hr, bool sigma,
                  gu_stats &s, int &dmax, hipStream_t st, int queue, bool allow_lds) {
  if (queue == 1 || (queue == 0 && nsrc > kPsCap)) return run_persource_queue(g, sources, nsrc, hr, sigma, s, dmax, st);
  const bool rl = g->LR != nullptr;
  int rc = ensure_pws(g, rl || hr.dist == nullptr);
  if (rc) return rc;
  PsWorkspace &p = g->pws;
  const size_t n = (size_t)g->n;
  if (rl && sigma && (rc = ensure_sscratch(g))) return rc;
  const size_t srow = (g->LE ? 2 : 1) * n;  // words a path-count scratch row takes
  const int32_t *nof = rl ? nof_host(g).data() : nullptr;
  u64 *part = g->part_h;
  int *depth = g->misc_h + 64;
  int ovf_any = 0;
  for (int b0 = 0; b0 < nsrc; b0 += kPsCap) {
    const int nb = std::min(kPsCap, nsrc - b0);
    void **t = g->rowtab_h;
    for (int i = 0; i < nb; i++) {
      g->srcs_h[i] = rl ? nof[sources[b0 + i]] : sources[b0 + i];
      int *dr = hr.dist ? hr.dist[b0 + i] : nullptr;
      u64 *sr = sigma && hr.sigma ? hr.sigma[b0 + i] : nullptr;
      if (sigma && !sr) return GU_ERR_INVALID;  // path counts need every row
      if (rl) {
        t[i] = p.dscratch + (size_t)i * n;
        t[256 + i] = sigma ? (void *)(p.sscratch + (size_t)i * srow) : nullptr;
        t[1024 + i] = dr;
        t[1280 + i] = sr;
      } else {
        // a source whose distances are not wanted works in the scratch row
        if (!dr) {
          rc = ensure_pws(g, true);
          if (rc) return rc;
        }
        t[i] = dr ? (void *)dr : (void *)(p.dscratch + (size_t)i * n);
        t[256 + i] = sr;
      }
    }
    GU_HIP(hipMemcpyAsync(p.srcs, g->srcs_h, sizeof(int) * nb, hipMemcpyHostToDevice, st));
    GU_HIP(hipMemcpyAsync(g->rowtab, t, sizeof(void *) * kRowTab, hipMemcpyHostToDevice, st));
    if (b0 == 0) k_fill_i32<<<1, 64, 0, st>>>(p.ovf, 1, 0);
    const int4 *LE = rl ? g->LE : nullptr;
    PsArgs a{rl ? g->LR : g->R, rl ? g->LC : g->C, LE, g->n, p.srcs, (int *const *)g->rowtab,
             sigma ? (u64 *const *)(g->rowtab + 256) : nullptr, p.queues, p.depth, p.part, p.ovf};
    a.q.nsrc = nb;
    auto fetch = [&]() -> int {
      GU_HIP(hipMemcpyAsync(part, p.part, sizeof(u64) * 4 * nb, hipMemcpyDeviceToHost, st));
      GU_HIP(hipMemcpyAsync(depth, p.depth, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
      GU_HIP(hipMemcpyAsync(g->misc_h, p.ovf, sizeof(int), hipMemcpyDeviceToHost, st));
      GU_HIP(hipStreamSynchronize(st));
      return GU_OK;
    };
    // undirected degree <= 4 graphs, with path counts: the LDS level-set
    // search (rows through the distance row); a launch in which a level set
    // overflowed runs again on the distance-reading kernel.  Searches
    // without path counts stay on that kernel: 64-source diameter sampling
    // on a 2048^2 grid took 34.4 ms there against 39.9 ms over level sets.
    bool lds = allow_lds && lds_ok(g, LE, sigma);
    const bool rows = hr.dist != nullptr || sigma;
    if (lds && sigma && (rc = ensure_slog(g))) return rc;
    if (lds) {
      PsLdsArgs c{LE,      g->n,   p.srcs, rows ? (int *const *)g->rowtab : nullptr,
                  sigma ? (u64 *const *)(g->rowtab + 256) : nullptr,
                  p.queues, p.depth, p.part, p.ovf, g->lds_flags, g->LE8};
      c.q.nsrc = nb;
      c.slog = p.slog;
      GuTraceRange tr("gu:per-source level sets");
      launch_ps_lds(c, sigma, rows, g->LE8 != nullptr, nb, st);
      GU_HIP(hipGetLastError());
      if ((rc = fetch())) return rc;
      for (int i = 0; i < nb; i++) lds &= (part[4 * i + 3] & 2) == 0;
      if (!lds) k_fill_i32<<<1, 64, 0, st>>>(p.ovf, 1, 0);
    }
    if (!lds) {
      GuTraceRange tr("gu:per-source search");
      // undirected graphs pull path counts from the previous level
      if (LE) {
        if (sigma && !g->directed)
          k_persource<true, true, true><<<nb, 1024, 0, st>>>(a);
        else if (sigma)
          k_persource<true, false, true><<<nb, 1024, 0, st>>>(a);
        else
          k_persource<false, false, true><<<nb, 1024, 0, st>>>(a);
      } else if (sigma && !g->directed) {
        k_persource<true, true><<<nb, 1024, 0, st>>>(a);
      } else if (sigma) {
        k_persource<true, false><<<nb, 1024, 0, st>>>(a);
      } else {
        k_persource<false, false><<<nb, 1024, 0, st>>>(a);
      }
      GU_HIP(hipGetLastError());
    }
    if (rl) {
      GuTraceRange tr("gu:per-source write-back");
      const dim3 ug((unsigned)std::max(1, std::min(pick_grid(g, 1), (g->n + 255) / 256)), (unsigned)nb);
      if (lds && rows) {  // 16-bit distances; path counts written for reached vertices only
        // chunks of about 4096 vertices (16 a thread), a multiple of 8 per row
        const unsigned wg = (unsigned)std::max<long long>(8, ((g->n + 4095) / 4096 + 7) / 8 * 8);
        if (sigma)  // records (distance and path count) of the search with path counts
          k_ps_unpermute_rec<<<dim3(wg, (unsigned)nb), 256, 0, st>>>(
              (const ps_rec_t *const *)(g->rowtab + 256), hr.dist ? (int *const *)(g->rowtab + 1024) : nullptr,
              (u64 *const *)(g->rowtab + 1280), g->nof, g->n);
        else
          k_ps_unpermute_ds<<<dim3(wg, (unsigned)nb), 256, 0, st>>>(
              (int *const *)g->rowtab, nullptr, hr.dist ? (int *const *)(g->rowtab + 1024) : nullptr, nullptr,
              g->nof, g->n);
      }
      if (hr.dist && !lds)
        k_ps_unpermute<int><<<ug, 256, 0, st>>>((int *const *)g->rowtab, (int *const *)(g->rowtab + 1024),
                                                g->nof, g->n);
      if (sigma && !lds)
        k_ps_unpermute<u64><<<ug, 256, 0, st>>>((u64 *const *)(g->rowtab + 256),
                                                (u64 *const *)(g->rowtab + 1280), g->nof, g->n);
      GU_HIP(hipGetLastError());
    }
    if (!lds && (rc = fetch())) return rc;
    ovf_any |= g->misc_h[0];
    for (int i = 0; i < nb; i++) {
      s.teps_units += (int64_t)part[4 * i];
      s.reached += (int64_t)part[4 * i + 1];
      s.dag_edges += (int64_t)part[4 * i + 2];
      dmax = std::max(dmax, depth[i]);
      if (hr.ecc) hr.ecc[b0 + i] = depth[i];
      if (hr.wrap) hr.wrap[b0 + i] = sigma && (part[4 * i + 3] & 1) != 0;
    }
    s.batches++;
  }
  s.sigma_overflow |= ovf_any;
  return GU_OK;
}

// Brandes on the per-source engine: forward and backward of each source in
// one workgroup (k_ps_bc), then the FP64 accumulation into the targets in
// source order (k_ps_bc_reduce).
// Brandes scratch rows, kPsCap of each: (1 + dependency) / path count of
// each vertex (qscratch) and, when the dependencies are not written into the caller's
// rows directly, the dependencies (xscratch).
int ensure_bc_scratch(gu_graph *g, bool need_x) {
  PsWorkspace &p = g->pws;
  const size_t n = (size_t)g->n;
  if (!p.qscratch) {
    GU_HIP(dalloc(&p.qscratch, (size_t)kPsCap * n));
    p.bytes += (size_t)kPsCap * n * 4;
  }
  if (need_x && !p.xscratch) {
    GU_HIP(dalloc(&p.xscratch, (size_t)kPsCap * n));
    p.bytes += (size_t)kPsCap * n * 4;
  }
  return GU_OK;
}

int run_persource_bc(gu_graph *g, const int *sources, int nsrc, const HostRows &hr, gu_stats &s,
                     int &dmax, hipStream_t st) {
  int rc = ensure_pws(g, true);
  if (rc) return rc;
  PsWorkspace &p = g->pws;
  const size_t n = (size_t)g->n;
  if ((rc = ensure_sscratch(g))) return rc;
  const bool rl = g->LR != nullptr;
  bool need_x = rl || hr.delta == nullptr;
  for (int i = 0; !need_x && i < nsrc; i++) need_x = hr.delta[i] == nullptr;
  if ((rc = ensure_bc_scratch(g, need_x))) return rc;
  u64 *part = g->part_h;
  int *depth = g->misc_h + 64;
  int ovf_any = 0;
  for (int b0 = 0; b0 < nsrc; b0 += kPsCap) {
    const int nb = std::min(kPsCap, nsrc - b0);
    const int32_t *nof = rl ? nof_host(g).data() : nullptr;
    void **t = g->rowtab_h;
    for (int i = 0; i < nb; i++) {
      g->srcs_h[i] = rl ? nof[sources[b0 + i]] : sources[b0 + i];
      t[i] = p.dscratch + (size_t)i * n;
      t[256 + i] = p.sscratch + (size_t)i * n;
      float *x = hr.delta ? hr.delta[b0 + i] : nullptr;
      // in cluster order the dependencies live in scratch rows (new ids)
      t[512 + i] = (x && !rl) ? (void *)x : (void *)(p.xscratch + (size_t)i * n);
      t[768 + i] = hr.bc ? (void *)hr.bc[b0 + i] : nullptr;
      t[1024 + i] = x;
      t[1280 + i] = p.qscratch + (size_t)i * n;
    }
    GU_HIP(hipMemcpyAsync(p.srcs, g->srcs_h, sizeof(int) * nb, hipMemcpyHostToDevice, st));
    GU_HIP(hipMemcpyAsync(g->rowtab, t, sizeof(void *) * kRowTab, hipMemcpyHostToDevice, st));
    if (b0 == 0) k_fill_i32<<<1, 64, 0, st>>>(p.ovf, 1, 0);
    const int4 *LE = rl ? g->LE : nullptr;
    PsBcArgs a{rl ? g->LR : g->R, rl ? g->LC : g->C, LE, g->n, p.srcs, (int *const *)g->rowtab,
               (u64 *const *)(g->rowtab + 256), (float *const *)(g->rowtab + 512),
               (float *const *)(g->rowtab + 1280), p.queues, p.depth, p.part, p.ovf, g->lds_flags, g->LE8};
    // undirected degree <= 4 graphs: the LDS level-set forward; a launch in
    // which a level set overflowed runs again on k_ps_bc
    bool lds = lds_ok(g, LE, true);
    if (lds && (rc = ensure_slog(g))) return rc;
    a.slog = p.slog;
    if (lds) {
      GuTraceRange tr("gu:per-source brandes level sets");
      if (g->LE8)
        k_ps_bc_lds<true><<<nb, 1024, 0, st>>>(a);
      else
        k_ps_bc_lds<false><<<nb, 1024, 0, st>>>(a);
      GU_HIP(hipGetLastError());
      GU_HIP(hipMemcpyAsync(part, p.part, sizeof(u64) * 4 * nb, hipMemcpyDeviceToHost, st));
      GU_HIP(hipStreamSynchronize(st));
      for (int i = 0; i < nb; i++) lds &= (part[4 * i + 3] & 2) == 0;
      if (!lds) k_fill_i32<<<1, 64, 0, st>>>(p.ovf, 1, 0);
    }
    if (!lds) {
      GuTraceRange tr("gu:per-source brandes");
      if (LE && g->directed)
        k_ps_bc<false, true><<<nb, 1024, 0, st>>>(a);
      else if (LE)
        k_ps_bc<true, true><<<nb, 1024, 0, st>>>(a);
      else if (g->directed)
        k_ps_bc<false><<<nb, 1024, 0, st>>>(a);
      else
        k_ps_bc<true><<<nb, 1024, 0, st>>>(a);
      GU_HIP(hipGetLastError());
    }
    if (rl && hr.delta) {
      const dim3 ug((unsigned)std::max(1, std::min(pick_grid(g, 1), (g->n + 255) / 256)), (unsigned)nb);
      k_ps_unpermute<float><<<ug, 256, 0, st>>>((float *const *)(g->rowtab + 512),
                                                (float *const *)(g->rowtab + 1024), g->nof, g->n);
    }
    GuTraceRange tr_red("gu:brandes reduce");
    if (hr.bc)
      k_ps_bc_reduce<<<pick_grid(g, 8), 256, 0, st>>>((float *const *)(g->rowtab + 512),
                                                       (double *const *)(g->rowtab + 768), nb, g->n,
                                                       rl ? g->nof : nullptr);
    GU_HIP(hipGetLastError());
    GU_HIP(hipMemcpyAsync(part, p.part, sizeof(u64) * 4 * nb, hipMemcpyDeviceToHost, st));
    GU_HIP(hipMemcpyAsync(depth, p.depth, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    GU_HIP(hipMemcpyAsync(g->misc_h, p.ovf, sizeof(int), hipMemcpyDeviceToHost, st));
    GU_HIP(hipStreamSynchronize(st));
    ovf_any |= g->misc_h[0];
    for (int i = 0; i < nb; i++) {
      s.teps_units += (int64_t)part[4 * i];
      s.reached += (int64_t)part[4 * i + 1];
      s.dag_edges += (int64_t)part[4 * i + 2];
      dmax = std::max(dmax, depth[i]);
      if (hr.wrap) hr.wrap[b0 + i] = (part[4 * i + 3] & 1) != 0;
    }
    s.batches++;
  }
  s.sigma_overflow |= ovf_any;
  return GU_OK;
}

// Engine choice from the degree profile: bounded-degree, sparse graphs are
// the high-diameter ones whose sources do not share levels.
int resolve_engine(gu_graph *g, const gu_config &cfg, int job) {
  (void)job;
  if (cfg.engine == GU_ENGINE_BATCHED || cfg.engine == GU_ENGINE_PER_SOURCE) return cfg.engine;
  return (g->max_deg <= 64 && g->avg_deg <= 8.0) ? GU_ENGINE_PER_SOURCE : GU_ENGINE_BATCHED;
}

int resolve_variant(gu_graph *g, const gu_config &cfg) {
  int v = cfg.variant;
  if (v >= V_LINEAR && v <= V_SHUFFLE) return v;
  if (g->tuned_variant >= 0) return g->tuned_variant;
  // Degree-profile default: heavy-tailed graphs need hub adjacency lists
  // split across a workgroup (the scan expander: 482 GTEPS on R-MAT s20
  // against 470 for the merge-path one, which also pays three scan launches
  // per level); bounded-degree graphs do well with one wave per frontier
  // vertex.
  return g->max_deg > 4096 ? V_SCAN : V_SHUFFLE;
}

int resolve_width(const gu_config &cfg, int nsrc) {
  int w = cfg.batch_width;
  if (w == 64 || w == 128 || w == 256) return w;
  if (w != 0) return -1;
  return nsrc <= 64 ? 64 : (nsrc <= 128 ? 128 : 256);
}

// The launch shape of a job of nsrc sources: fields left at "auto" take the
// autotuned shape of the job family, for launches of about the size it was
// tuned on or larger (small launches keep the degree-profile defaults).
gu_config job_config(const gu_graph *g, gu_config cfg, Job job, int nsrc) {
  const gu_graph::Tuned &t = g->tuned[job == JOB_BC ? 1 : 0];
  if (2 * nsrc < t.sources) return cfg;
  if (cfg.engine == GU_ENGINE_AUTO) cfg.engine = t.engine;
  if (cfg.variant == GU_VARIANT_AUTO && t.variant >= 0) cfg.variant = t.variant;
  if (cfg.wg_size == 0) cfg.wg_size = t.wg;
  if (cfg.alpha == 0) cfg.alpha = t.alpha;
  if (cfg.beta == 0) cfg.beta = t.beta;
  if (cfg.batch_width == 0 && t.width > 0) {
    const int w = nsrc <= 64 ? 64 : (nsrc <= 128 ? 128 : 256);
    cfg.batch_width = std::min(w, t.width);
  }
  return cfg;
}

// Allocates every workspace buffer that a launch of 1 to kPsCap sources of
// each job (search with or without path counts, Brandes, diameter) can need
// on this handle with this configuration, so that such launches allocate
// nothing.  The batched Brandes stack is sized for 24n entries, room for a
// chunk of 8 levels on top of 14n entries already pushed.
int reserve_launches(gu_graph *g, const gu_config &base) {
  GU_HIP(hipSetDevice(g->device));
  const size_t n = (size_t)g->n;
  const bool rl = g->LR != nullptr;
  const int4 *LE = rl ? g->LE : nullptr;
  int rc = GU_OK;
  // Brandes and the full width first: the batched workspace is sized for the
  // widest batch, and a narrower one uses its front
  for (Job job : {JOB_BC, JOB_BFS, JOB_DIAM})
    for (int nsrc : {kPsCap, 1}) {
      const gu_config cfg = job_config(g, base, job, nsrc);
      if (cfg.engine < 0 || cfg.engine > GU_ENGINE_PER_SOURCE) return GU_ERR_INVALID;
      if (resolve_engine(g, cfg, job) == GU_ENGINE_PER_SOURCE) {
        const bool sigma = job != JOB_DIAM;  // searches may ask for path counts
        if ((rc = ensure_pws(g, true))) return rc;
        if (sigma && (job == JOB_BC || rl) && (rc = ensure_sscratch(g))) return rc;
        if (job == JOB_BC && (rc = ensure_bc_scratch(g, true))) return rc;
        if (sigma && lds_ok(g, LE, true) && (rc = ensure_slog(g))) return rc;
        if (cfg.source_queue == 1 && (rc = ensure_jobs(g, kPsCap))) return rc;
      } else {
        const int width = resolve_width(cfg, nsrc);
        if (width < 0) return GU_ERR_INVALID;
        const bool stack = job == JOB_BC;
        if ((rc = ensure_ws(g, width / 64, stack, stack ? 24 * n : 0, nullptr))) return rc;
      }
    }
  return GU_OK;
}

int run_job(gu_graph *g, const int32_t *sources, int32_t nsrc, Job job, const HostRows &hr,
            int32_t *diam, const gu_config *cfgp, gu_stats *stats) {
  if (!g) return GU_ERR_INVALID;
  gu_config cfg{};
  if (cfgp) cfg = *cfgp;
  cfg = job_config(g, cfg, job, nsrc);
  if (cfg.direction < 0 || cfg.direction > 2) return GU_ERR_INVALID;
  if (cfg.variant < 0 || cfg.variant > GU_VARIANT_AUTO) return GU_ERR_INVALID;
  if (cfg.engine < 0 || cfg.engine > GU_ENGINE_PER_SOURCE) return GU_ERR_INVALID;
  if (cfg.source_queue < 0 || cfg.source_queue > 2) return GU_ERR_INVALID;
  if (cfg.wg_size != 0 && cfg.wg_size != 256 && cfg.wg_size != 512 && cfg.wg_size != 1024)
    return GU_ERR_INVALID;
  if (nsrc < 0 || (nsrc > 0 && !sources)) return GU_ERR_INVALID;
  for (int i = 0; i < nsrc; i++)
    if (sources[i] < 0 || sources[i] >= g->n) return GU_ERR_RANGE;
  const int width = resolve_width(cfg, nsrc);
  if (width < 0) return GU_ERR_INVALID;
  const int variant = resolve_variant(g, cfg);
  // Path counts are computed for Brandes and when the caller asks for them.
  const bool want_sigma = job == JOB_BC || (job == JOB_BFS && hr.sigma != nullptr);
  if (hr.wrap) std::memset(hr.wrap, 0, sizeof(int32_t) * (size_t)nsrc);
  hipStream_t st = (hipStream_t)cfg.stream;
  GU_HIP(hipSetDevice(g->device));
  EventPair ev;
  GU_HIP(hipEventCreate(&ev.a));
  GU_HIP(hipEventCreate(&ev.b));
  GU_HIP(hipEventRecord(ev.a, st));
  gu_stats s{};
  s.variant_used = variant;
  s.batch_width_used = width;
  s.engine_used = resolve_engine(g, cfg, job);
  int rc = GU_OK;
  int dmax = 0;
  if (s.engine_used == GU_ENGINE_PER_SOURCE && nsrc > 0)
    rc = job == JOB_BC ? run_persource_bc(g, sources, nsrc, hr, s, dmax, st)
                       : run_persource(g, sources, nsrc, hr, want_sigma, s, dmax, st, cfg.source_queue);
  for (int b0 = 0; s.engine_used == GU_ENGINE_BATCHED && b0 < nsrc && rc == GU_OK; b0 += width) {
    int nb = std::min(width, nsrc - b0);
    void **t = g->rowtab_h;
    for (int i = 0; i < nb; i++) {
      t[i] = hr.dist ? (void *)hr.dist[b0 + i] : nullptr;
      t[256 + i] = hr.sigma ? (void *)hr.sigma[b0 + i] : nullptr;
      t[512 + i] = hr.delta ? (void *)hr.delta[b0 + i] : nullptr;
      t[768 + i] = hr.bc ? (void *)hr.bc[b0 + i] : nullptr;
    }
    GU_HIP(hipMemcpyAsync(g->rowtab, t, sizeof(void *) * kRowTab, hipMemcpyHostToDevice, st));
    BatchOut o;
    if (hr.dist) o.dist = (int *const *)g->rowtab;
    if (hr.sigma) o.sigma = (u64 *const *)(g->rowtab + 256);
    if (hr.delta) o.delta = (float *const *)(g->rowtab + 512);
    if (hr.bc) o.bc = (double *const *)(g->rowtab + 768);
    if (hr.ecc) o.ecc = g->ecc;
    RunInfo info;
    if (width == 64)
      rc = run_batch<1>(g, sources + b0, nb, job, want_sigma, cfg, variant, o, &info, st);
    else if (width == 128)
      rc = run_batch<2>(g, sources + b0, nb, job, want_sigma, cfg, variant, o, &info, st);
    else
      rc = run_batch<4>(g, sources + b0, nb, job, want_sigma, cfg, variant, o, &info, st);
    if (rc) break;
    if (hr.ecc) std::memcpy(hr.ecc + b0, g->ecc_h, sizeof(int) * nb);
    if (hr.wrap)
      for (int i = 0; i < nb; i++) hr.wrap[b0 + i] = want_sigma && info.overflow;
    s.batches++;
    s.topdown_levels += info.td;
    s.bottomup_levels += info.bu;
    s.sigma_overflow |= info.overflow;
    s.teps_units += (int64_t)info.teps;
    s.reached += (int64_t)info.reached;
    s.dag_edges += (int64_t)info.dag;
    dmax = std::max(dmax, info.last_level);
  }
  GU_HIP(hipEventRecord(ev.b, st));
  GU_HIP(hipEventSynchronize(ev.b));
  float ms = 0;
  GU_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
  s.device_ms = ms;
  s.max_level = dmax;
  if (diam) *diam = dmax;
  if (stats) *stats = s;
  return rc;
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

const char *gu_strerror(int s) {
  switch (s) {
    case GU_OK: return "ok";
    case GU_ERR_INVALID: return "invalid argument";
    case GU_ERR_HIP: return g_last_err[0] ? g_last_err : "HIP runtime error";
    case GU_ERR_NOMEM: return "out of memory";
    case GU_ERR_RANGE: return "vertex id out of range";
    case GU_ERR_PARSE: return "malformed graph file";
    case GU_ERR_WEIGHTED: return "weighted graphs are not supported";
    case GU_ERR_IO: return "cannot open file";
    case GU_ERR_NODEVICE: return "no HIP device";
    case GU_ERR_CAPACITY: return "output capacity exceeded";
    case GU_ERR_TIMEOUT: return "timed out";
    case GU_ERR_UNSUPPORTED: return "not supported on this host";
    default: return "unknown status";
  }
}

int gu_device_count(int32_t *count) {
  if (!count) return GU_ERR_INVALID;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return GU_ERR_NODEVICE;
  }
  *count = c;
  return GU_OK;
}

int gu_choose_device(int32_t requested, gu_device_props *props) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GU_ERR_NODEVICE;
  int dev = requested;
  if (requested < 0) {
    double best = -1;
    for (int i = 0; i < count; i++) {
      hipDeviceProp_t p;
      GU_HIP(hipGetDeviceProperties(&p, i));
      double bw = (p.memoryClockRate * 1000.0) * (p.memoryBusWidth / 8 * 2) / 1.0e9;
      if (bw > best) {
        best = bw;
        dev = i;
      }
    }
  } else if (requested >= count) {
    return GU_ERR_INVALID;
  }
  GU_HIP(hipSetDevice(dev));
  if (props) {
    hipDeviceProp_t p;
    GU_HIP(hipGetDeviceProperties(&p, dev));
    std::memset(props, 0, sizeof(*props));
    std::snprintf(props->name, sizeof(props->name), "%s", p.name);
    std::snprintf(props->arch, sizeof(props->arch), "%s", p.gcnArchName);
    props->device = dev;
    props->compute_units = p.multiProcessorCount;
    props->max_threads_per_block = p.maxThreadsPerBlock;
    props->wave_size = p.warpSize;
    props->total_mem = (int64_t)p.totalGlobalMem;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) props->free_mem = (int64_t)fr;
    props->mem_bandwidth_gbs = (p.memoryClockRate * 1000.0) * (p.memoryBusWidth / 8 * 2) / 1.0e9;
    props->l2_bytes = p.l2CacheSize;
    props->clock_khz = p.clockRate;
  }
  return GU_OK;
}

struct gu_timer_s {
  hipEvent_t a, b;
  hipStream_t s;
};

int gu_timer_start(void *stream, void **timer) {
  if (!timer) return GU_ERR_INVALID;
  *timer = nullptr;
  gu_timer_s *t = new (std::nothrow) gu_timer_s{};
  if (!t) return GU_ERR_NOMEM;
  t->s = (hipStream_t)stream;
  hipError_t e = hipEventCreate(&t->a);
  if (e == hipSuccess) e = hipEventCreate(&t->b);
  if (e == hipSuccess) e = hipEventRecord(t->a, t->s);
  if (e != hipSuccess) {
    if (t->a) (void)hipEventDestroy(t->a);
    if (t->b) (void)hipEventDestroy(t->b);
    delete t;
    return fail(e);
  }
  *timer = t;
  return GU_OK;
}

int gu_timer_stop(void *timer, double *ms) {
  if (!timer) return GU_ERR_INVALID;
  gu_timer_s *t = (gu_timer_s *)timer;
  float f = 0;
  hipError_t e = hipEventRecord(t->b, t->s);
  if (e == hipSuccess) e = hipEventSynchronize(t->b);
  if (e == hipSuccess) e = hipEventElapsedTime(&f, t->a, t->b);
  (void)hipEventDestroy(t->a);
  (void)hipEventDestroy(t->b);
  delete t;  // the timer is consumed either way
  if (e != hipSuccess) return fail(e);
  if (ms) *ms = f;
  return GU_OK;
}

int gu_malloc(void **p, size_t bytes) {
  if (!p) return GU_ERR_INVALID;
  GU_HIP(hipMalloc(p, bytes ? bytes : 1));
  return GU_OK;
}
int gu_free(void *p) {
  GU_HIP(hipFree(p));
  return GU_OK;
}
int gu_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  GU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GU_OK;
}
int gu_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  GU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GU_OK;
}
int gu_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  GU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GU_OK;
}
int gu_memset_d(void *dst, int value, size_t bytes) {
  GU_HIP(hipMemset(dst, value, bytes));
  return GU_OK;
}
int gu_synchronize(void) {
  GU_HIP(hipDeviceSynchronize());
  return GU_OK;
}

}  // extern "C"

namespace {
// Locality order for the per-source engine.  On a bounded-degree, high-
// diameter graph each source's frontier is a thin ring; with the input
// numbering a visited vertex touches separate cache lines for itself and
// for each neighbour, and the ring returns to them only after they were
// evicted.  Grouping vertices into clusters of 32 grown by BFS (from the
// lowest-numbered unassigned vertex, over unassigned vertices) and giving
// each cluster consecutive ids puts most neighbours in the same 128-byte
// line of every per-vertex array.  Adjacency order within a row is kept.
void cluster_order(int n, const int32_t *R, const int32_t *C, std::vector<int32_t> &nof,
                   std::vector<int32_t> &oon) {
  nof.assign((size_t)n, -1);
  oon.assign((size_t)n, 0);
  std::vector<int32_t> q;
  int next = 0;
  for (int s = 0; s < n; s++) {
    if (nof[(size_t)s] >= 0) continue;
    q.clear();
    q.push_back(s);
    int got = 0;
    for (size_t h = 0; h < q.size() && got < 32; h++) {
      const int v = q[h];
      if (nof[(size_t)v] >= 0) continue;
      nof[(size_t)v] = next;
      oon[(size_t)next] = v;
      next++;
      got++;
      for (int k = R[v]; k < R[v + 1]; k++)
        if (nof[(size_t)C[k]] < 0) q.push_back(C[k]);
    }
  }
}

// Bipartite and connected (kLds* bits) for an undirected graph, by one BFS
// 2-colouring over every component: O(n + m) on the host.
int lds_graph_facts(int n, const int32_t *R, const int32_t *C) {
  std::vector<int8_t> col((size_t)n, -1);
  std::vector<int32_t> q((size_t)n);
  bool bip = true;
  int comps = 0;
  for (int s = 0; s < n; s++) {
    if (col[(size_t)s] >= 0) continue;
    comps++;
    size_t h = 0, t = 0;
    q[t++] = s;
    col[(size_t)s] = 0;
    while (h < t) {
      const int v = q[h++];
      for (int k = R[v]; k < R[v + 1]; k++) {
        const int w = C[k];
        if (col[(size_t)w] < 0) {
          col[(size_t)w] = (int8_t)(col[(size_t)v] ^ 1);
          q[t++] = w;
        } else if (col[(size_t)w] == col[(size_t)v]) {
          bip = false;
        }
      }
    }
  }
  return (bip ? gu::kLdsBipartite : 0) | (comps == 1 ? gu::kLdsConnected : 0);
}

int build_cluster_copy(gu_graph *g, const int32_t *R, const int32_t *C) {
  const int n = g->n;
  std::vector<int32_t> oon, LR((size_t)n + 1), LC((size_t)std::max<long long>(g->m, 1));
  cluster_order(n, R, C, g->nof_h, oon);
  LR[0] = 0;
  for (int u = 0; u < n; u++) {
    const int v = oon[(size_t)u];
    int o = LR[(size_t)u];
    for (int k = R[v]; k < R[v + 1]; k++) LC[(size_t)o++] = g->nof_h[(size_t)C[k]];
    LR[(size_t)u + 1] = o;
  }
  if (dalloc(&g->LR, (size_t)n + 1) != hipSuccess || dalloc(&g->LC, LC.size()) != hipSuccess ||
      dalloc(&g->nof, (size_t)std::max(n, 1)) != hipSuccess)
    return GU_ERR_NOMEM;
  if (g->max_deg <= 4 && n > 0) {
    if (!g->directed) g->lds_flags = lds_graph_facts(n, R, C);
    std::vector<int4> LE((size_t)n);
    for (int u = 0; u < n; u++) {
      int q[4] = {-1, -1, -1, -1};
      for (int k = LR[(size_t)u]; k < LR[(size_t)u + 1]; k++) q[k - LR[(size_t)u]] = LC[(size_t)k];
      LE[(size_t)u] = make_int4(q[0], q[1], q[2], q[3]);
    }
    if (hipMalloc((void **)&g->LE, sizeof(int4) * (size_t)n) != hipSuccess) return GU_ERR_NOMEM;
    if (hipMemcpy(g->LE, LE.data(), sizeof(int4) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
      return GU_ERR_HIP;
    std::vector<u64> LE8((size_t)n);
    bool fits = true;
    for (int u = 0; u < n && fits; u++) {
      const int q[4] = {LE[(size_t)u].x, LE[(size_t)u].y, LE[(size_t)u].z, LE[(size_t)u].w};
      u64 p = 0;
      for (int j = 0; j < 4; j++) {
        const long long dlt = q[j] < 0 ? (long long)gu::kLe8None : (long long)q[j] - u;
        fits &= q[j] < 0 || (dlt > gu::kLe8None && dlt <= 32767);
        p |= (u64)(uint16_t)(int16_t)dlt << (16 * j);
      }
      LE8[(size_t)u] = p;
    }
    if (fits) {
      if (hipMalloc((void **)&g->LE8, sizeof(u64) * (size_t)n) != hipSuccess) return GU_ERR_NOMEM;
      if (hipMemcpy(g->LE8, LE8.data(), sizeof(u64) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return GU_ERR_HIP;
    }
  }
  if (hipMemcpy(g->LR, LR.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->LC, LC.data(), sizeof(int) * LC.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (n > 0 && hipMemcpy(g->nof, g->nof_h.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess))
    return GU_ERR_HIP;
  return GU_OK;
}

// Pinned staging and the launch tables every handle owns.
int alloc_handle_bufs(gu_graph *g) {
  if (hipHostMalloc((void **)&g->ctrl_h, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&g->srcs_h, sizeof(int) * 256, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&g->rowtab_h, sizeof(void *) * kRowTab, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&g->ecc_h, sizeof(int) * 256, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&g->misc_h, sizeof(int) * 320, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&g->part_h, sizeof(u64) * 4 * 256, hipHostMallocDefault) != hipSuccess)
    return GU_ERR_NOMEM;
  std::memset(g->ctrl_h, 0, sizeof(Ctrl));
  std::memset(g->rowtab_h, 0, sizeof(void *) * kRowTab);
  if (dalloc(&g->rowtab, kRowTab) != hipSuccess || dalloc(&g->ecc, 256) != hipSuccess)
    return GU_ERR_NOMEM;
  return GU_OK;
}
}  // namespace

extern "C" {

int gu_graph_create(const int32_t *R, const int32_t *C, const int32_t *F, int32_t n, int64_t m,
                    int32_t directed, gu_graph_t *out) {
  if (!out || !R || n < 0 || m < 0 || m >= INT32_MAX || (m > 0 && !C)) return GU_ERR_INVALID;
  *out = nullptr;
  // host-side checks of the CSR shape (the reference trusts its parser),
  // before any device call
  if (R[0] != 0 || (int64_t)R[n] != m) return GU_ERR_INVALID;
  int maxd = 0;
  long long nonzero = 0;
  for (int32_t v = 0; v < n; v++) {
    if (R[v + 1] < R[v]) return GU_ERR_INVALID;
    maxd = std::max(maxd, R[v + 1] - R[v]);
    nonzero += R[v + 1] > R[v];
  }
  bool strict = true;  // rows strictly increasing: no parallel edges
  // An undirected graph's entries are a symmetric multiset: (v, w) occurs as
  // often as (w, v).  Then the sum over entries of h(v, w) equals that of
  // h(w, v) for any h; with a 64-bit mixing hash an asymmetric CSR fails
  // the equality except with probability about 2^-64, and a symmetric one
  // always passes.
  u64 fwd = 0, bwd = 0;
  for (int32_t v = 0; v < n; v++)
    for (int64_t k = R[v]; k < R[v + 1]; k++) {
      if (C[k] < 0 || C[k] >= n) return GU_ERR_RANGE;
      strict &= k == R[v] || C[k] > C[k - 1];
      if (!directed) {
        fwd += mix64(((u64)(uint32_t)v << 32) | (uint32_t)C[k]);
        bwd += mix64(((u64)(uint32_t)C[k] << 32) | (uint32_t)v);
      }
    }
  if (fwd != bwd) return GU_ERR_INVALID;  // declared undirected, not symmetric
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GU_ERR_NODEVICE;
  gu_graph *g = new (std::nothrow) gu_graph;
  if (!g) return GU_ERR_NOMEM;
  g->n = n;
  g->m = m;
  g->directed = directed;
  g->max_deg = maxd;
  g->avg_deg = nonzero ? (double)m / (double)nonzero : 0.0;
  g->simple = strict;
  int rc = GU_OK;
  do {
    if (hipGetDevice(&g->device) != hipSuccess) { rc = GU_ERR_HIP; break; }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, g->device) != hipSuccess) { rc = GU_ERR_HIP; break; }
    g->cus = p.multiProcessorCount;
    if (dalloc(&g->R, (size_t)n + 1) != hipSuccess || dalloc(&g->C, (size_t)m) != hipSuccess ||
        dalloc(&g->F, (size_t)m) != hipSuccess || dalloc(&g->ctrl, 1) != hipSuccess) {
      rc = GU_ERR_NOMEM;
      break;
    }
    if ((rc = alloc_handle_bufs(g)) != GU_OK) break;
    if (hipMemcpy(g->R, R, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
        (m > 0 && hipMemcpy(g->C, C, sizeof(int) * (size_t)m, hipMemcpyHostToDevice) != hipSuccess)) {
      rc = GU_ERR_HIP;
      break;
    }
    if (directed) {
      std::vector<int32_t> RT, CT;
      transpose_csr(n, m, R, C, RT, CT);
      if (dalloc(&g->RT, (size_t)n + 1) != hipSuccess || dalloc(&g->CT, (size_t)m) != hipSuccess) {
        rc = GU_ERR_NOMEM;
        break;
      }
      if (hipMemcpy(g->RT, RT.data(), sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice) != hipSuccess ||
          (m > 0 && hipMemcpy(g->CT, CT.data(), sizeof(int) * (size_t)m, hipMemcpyHostToDevice) != hipSuccess)) {
        rc = GU_ERR_HIP;
        break;
      }
      if ((rc = build_units(n, RT.data(), &g->pu)) != GU_OK) break;
      if ((rc = build_units(n, R, &g->ou)) != GU_OK) break;
    } else {
      if ((rc = build_units(n, R, &g->pu)) != GU_OK) break;
      g->ou = g->pu;
      g->ou.owned = false;
    }
    // the per-source engine's locality copy, on graphs it is chosen for
    if (n > 0 && maxd <= 64 && g->avg_deg <= 8.0 && (rc = build_cluster_copy(g, R, C)) != GU_OK) break;
    if (F) {
      if (m > 0 && hipMemcpy(g->F, F, sizeof(int) * (size_t)m, hipMemcpyHostToDevice) != hipSuccess) {
        rc = GU_ERR_HIP;
        break;
      }
    } else if (n > 0) {
      k_build_F<<<std::max(1, g->cus * 4), 256>>>(g->R, n, g->F);
      if (hipDeviceSynchronize() != hipSuccess) { rc = GU_ERR_HIP; break; }
    }
  } while (0);
  if (rc != GU_OK) {
    gu_graph_destroy(g);
    return rc;
  }
  *out = g;
  return GU_OK;
}

int gu_graph_destroy(gu_graph_t g) {
  if (!g) return GU_ERR_INVALID;
  if (g->views.load() != 0) return GU_ERR_INVALID;  // shared handles still use it
  clear_level_graphs(g);
  if (g->cap_stream) (void)hipStreamDestroy(g->cap_stream);
  g->ws.release();
  g->pws.release();
  if (!g->root) {
    if (g->R) (void)hipFree(g->R);
    if (g->C) (void)hipFree(g->C);
    if (g->F) (void)hipFree(g->F);
    g->pu.release();
    g->ou.release();
    if (g->RT) (void)hipFree(g->RT);
    if (g->CT) (void)hipFree(g->CT);
    if (g->LR) (void)hipFree(g->LR);
    if (g->LE) (void)hipFree(g->LE);
    if (g->LE8) (void)hipFree(g->LE8);
    if (g->LC) (void)hipFree(g->LC);
    if (g->nof) (void)hipFree(g->nof);
  } else {
    g->root->views.fetch_sub(1);
  }
  if (g->ctrl) (void)hipFree(g->ctrl);
  if (g->rowtab) (void)hipFree(g->rowtab);
  if (g->ecc) (void)hipFree(g->ecc);
  void *pinned[] = {g->ctrl_h, g->srcs_h, g->rowtab_h, g->ecc_h, g->misc_h, g->part_h};
  for (void *p : pinned)
    if (p) (void)hipHostFree(p);
  delete g;
  return GU_OK;
}

int gu_graph_share(gu_graph_t g, gu_graph_t *out) {
  if (!g || !out) return GU_ERR_INVALID;
  *out = nullptr;
  gu_graph *root = g->root ? g->root : g;
  gu_graph *v = new (std::nothrow) gu_graph;
  if (!v) return GU_ERR_NOMEM;
  v->n = root->n;
  v->m = root->m;
  v->directed = root->directed;
  v->R = root->R;
  v->C = root->C;
  v->F = root->F;
  v->cus = root->cus;
  v->device = root->device;
  v->max_deg = root->max_deg;
  v->avg_deg = root->avg_deg;
  v->simple = root->simple;
  v->RT = root->RT;
  v->CT = root->CT;
  v->LR = root->LR;
  v->LE = root->LE;
  v->LE8 = root->LE8;
  v->lds_flags = root->lds_flags;
  v->LC = root->LC;
  v->nof = root->nof;
  v->pu = root->pu;
  v->pu.owned = false;
  v->ou = root->ou;
  v->ou.owned = false;
  v->tuned_variant = root->tuned_variant;
  v->tuned_direction_alpha = root->tuned_direction_alpha;
  v->tuned[0] = root->tuned[0];
  v->tuned[1] = root->tuned[1];
  v->root = root;
  root->views.fetch_add(1);
  int rc = GU_OK;
  if (hipSetDevice(root->device) != hipSuccess) rc = GU_ERR_HIP;
  if (rc == GU_OK && dalloc(&v->ctrl, 1) != hipSuccess) rc = GU_ERR_NOMEM;
  if (rc == GU_OK) rc = alloc_handle_bufs(v);
  if (rc != GU_OK) {
    gu_graph_destroy(v);
    return rc;
  }
  *out = v;
  return GU_OK;
}

int gu_graph_arrays(gu_graph_t g, const int32_t **R, const int32_t **C, const int32_t **F,
                    int32_t *n, int64_t *m) {
  if (!g) return GU_ERR_INVALID;
  if (R) *R = g->R;
  if (C) *C = g->C;
  if (F) *F = g->F;
  if (n) *n = g->n;
  if (m) *m = g->m;
  return GU_OK;
}

int gu_graph_device_bytes(gu_graph_t g, int64_t *bytes) {
  if (!g || !bytes) return GU_ERR_INVALID;
  *bytes = (int64_t)(((size_t)g->n + 1) * 4 + (size_t)g->m * 8 + g->ws.bytes + g->pws.bytes);
  if (g->LR) *bytes += (int64_t)(((size_t)g->n + 1) * 4 + (size_t)g->m * 4 + (size_t)g->n * 4);
  if (g->LE) *bytes += (int64_t)((size_t)g->n * 16);
  if (g->LE8) *bytes += (int64_t)((size_t)g->n * 8);
  return GU_OK;
}

}  // extern "C"

namespace {
// Scan tiles of the three-phase counts scan (kernels of 256 threads).
long long lbs_scan_tiles(long long S) {
  const long long TS = 256LL * kScanIPT;
  return (S + TS - 1) / TS;
}

// The launch sequence of a load-balanced search over counts[0, S): up to
// kLbsSmall segments one k_lbs_small launch, else the three-phase scan into
// tsum (lbs_scan_tiles(S) + 2 words, the total at tsum[nt + 1]) and, with a
// result, the tile starts tlo (cap / kLbsXT + 2 ints) and the expansion.
// Items beyond cap are not written (the kernels test W > cap).
hipError_t lbs_launch(const int32_t *counts, long long S, int32_t *scan, int32_t *result, long long cap, int cus,
                      long long *tsum, long long *total_d, int *tlo, hipStream_t st) {
  const int T = 256;
  const long long nt = lbs_scan_tiles(S);
  const int grid = std::max(1, cus * 8);
  if (S <= kLbsSmall) {
    // without a result buffer only the scan and the total are wanted.  Two
    // workgroups of 512 per CU: 2^26 items in 1805 Zipf segments took 50.1
    // us, against 52.0 with four of 256 and 50.2 with one of 1024 (2 and 8
    // workgroups of 256 per CU ran at 848 and 1141 against 1230 G items/s)
    k_lbs_small<<<result ? std::max(1, cus * 2) : 1, kLbsSmallT, 0, st>>>(counts, (int)S, scan,
                                                                    result ? cap : -1, result, total_d);
  } else {
    k_cscan_reduce<<<grid, T, 0, st>>>(counts, S, tsum);
    k_cscan_tiles<<<1, T, 0, st>>>(tsum, nt, total_d);
    k_cscan_apply<<<grid, T, 0, st>>>(counts, S, tsum, scan, total_d);
    if (result) {
      k_lbs_tiles<<<grid, 256, 0, st>>>(scan, S, total_d, cap, tlo);
      k_lbs_expand<<<std::max(1, cus * 4), 256, 0, st>>>(scan, S, total_d, cap, tlo, result);
    }
  }
  return hipGetLastError();
}

// Largest inputs of gu_lbs_noalloc, and the bytes of its scratch: the tile
// sums and total of the three-phase scan (lbs_scan_tiles(S) + 2 words) and
// the tile starts of the expansion (W / kLbsXT + 2 ints).
constexpr long long kNaMaxS = GU_LBS_NOALLOC_MAX_SEGMENTS;
constexpr long long kNaMaxW = GU_LBS_NOALLOC_MAX_ITEMS;
constexpr size_t kNaTsumBytes = (size_t)((kNaMaxS + 256LL * kScanIPT - 1) / (256LL * kScanIPT) + 2) * 8;
constexpr size_t kNaTloBytes = (size_t)(kNaMaxW / kLbsXT + 2) * 4;
constexpr size_t kNaScratchBytes = (kNaTsumBytes + kNaTloBytes + 255) / 256 * 256;

}  // namespace

// Scratch of gu_lbs_noalloc: one copy per GPU, held until the process ends.
// Host code must not pass this name to a kernel (on the host it is a
// stand-in, not a device address); it uses the address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void *g_na_scratch[kNaMaxDevices] = {};

// Device address of `scratch` on device dev (the current device), looked up
// on the first call on that device and kept.
int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= kNaMaxDevices) return GU_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_scratch[dev]) {
    void *ptr = nullptr;
    GU_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return GU_ERR_HIP;
    g_na_scratch[dev] = ptr;
  }
  *out = g_na_scratch[dev];
  return GU_OK;
}

// Row i of a contiguous [nsrc x n] buffer.
template <class T>
std::vector<T *> contiguous_rows(T *base, int32_t nsrc, int32_t n) {
  std::vector<T *> r;
  if (base && nsrc > 0) {
    r.resize((size_t)nsrc);
    for (int32_t i = 0; i < nsrc; i++) r[(size_t)i] = base + (size_t)i * (size_t)n;
  }
  return r;
}
}  // namespace

extern "C" {

int gu_multi_search(gu_graph_t g, const int32_t *sources, int32_t nsrc, int32_t *dist,
                    uint64_t *sigma, const gu_config *cfg, gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  std::vector<int32_t *> dr = contiguous_rows(dist, nsrc, g->n);
  std::vector<u64 *> sr = contiguous_rows((u64 *)sigma, nsrc, g->n);
  HostRows hr;
  if (dist) hr.dist = dr.data();
  if (sigma) hr.sigma = sr.data();
  return run_job(g, sources, nsrc, JOB_BFS, hr, nullptr, cfg, stats);
}

int gu_multi_search_rows(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                         int32_t *const *dist_rows, uint64_t *const *sigma_rows, int32_t *ecc,
                         const gu_config *cfg, gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  HostRows hr;
  hr.dist = dist_rows;
  hr.sigma = (u64 *const *)sigma_rows;
  hr.ecc = ecc;
  // eccentricities without rows: the diameter job (no transposes)
  const Job job = (!dist_rows && !sigma_rows && ecc) ? JOB_DIAM : JOB_BFS;
  return run_job(g, sources, nsrc, job, hr, nullptr, cfg, stats);
}

int gu_betweenness(gu_graph_t g, const int32_t *sources, int32_t nsrc, float *delta, double *bc,
                   const gu_config *cfg, gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  std::vector<float *> xr = contiguous_rows(delta, nsrc, g->n);
  std::vector<double *> bt;
  if (bc) bt.assign((size_t)std::max(nsrc, 0), bc);
  HostRows hr;
  if (delta) hr.delta = xr.data();
  if (bc) hr.bc = bt.data();
  return run_job(g, sources, nsrc, JOB_BC, hr, nullptr, cfg, stats);
}

int gu_betweenness_rows(gu_graph_t g, const int32_t *sources, int32_t nsrc,
                        float *const *delta_rows, double *const *bc_targets,
                        const gu_config *cfg, gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  HostRows hr;
  hr.delta = delta_rows;
  hr.bc = bc_targets;
  return run_job(g, sources, nsrc, JOB_BC, hr, nullptr, cfg, stats);
}

}  // extern "C"

// Row forms with per-source wrap flags, for the query server (C++ linkage,
// declared in gu_internal.hpp).
int gu_search_rows_wrap(gu_graph_t g, const int32_t *sources, int32_t nsrc, int32_t *const *dist_rows,
                        uint64_t *const *sigma_rows, int32_t *ecc, int32_t *wrap, const gu_config *cfg,
                        gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  HostRows hr;
  hr.dist = dist_rows;
  hr.sigma = (u64 *const *)sigma_rows;
  hr.ecc = ecc;
  hr.wrap = wrap;
  const Job job = (!dist_rows && !sigma_rows && ecc) ? JOB_DIAM : JOB_BFS;
  return run_job(g, sources, nsrc, job, hr, nullptr, cfg, stats);
}

int gu_reserve_launches(gu_graph_t g, const gu_config *cfg) {
  if (!g) return GU_ERR_INVALID;
  gu_config c{};
  if (cfg) c = *cfg;
  return reserve_launches(g, c);
}

int gu_bc_rows_wrap(gu_graph_t g, const int32_t *sources, int32_t nsrc, float *const *delta_rows,
                    double *const *bc_targets, int32_t *wrap, const gu_config *cfg, gu_stats *stats) {
  if (!g || nsrc < 0) return GU_ERR_INVALID;
  HostRows hr;
  hr.delta = delta_rows;
  hr.bc = bc_targets;
  hr.wrap = wrap;
  return run_job(g, sources, nsrc, JOB_BC, hr, nullptr, cfg, stats);
}

extern "C" {

int gu_diameter(gu_graph_t g, const int32_t *sources, int32_t nsrc, int32_t *diameter,
                const gu_config *cfg, gu_stats *stats) {
  if (!diameter) return GU_ERR_INVALID;
  *diameter = 0;
  return run_job(g, sources, nsrc, JOB_DIAM, HostRows(), diameter, cfg, stats);
}

int gu_autotune(gu_graph_t g, const int32_t *sources, int32_t nsrc, int32_t kind,
                gu_tune_result *out) {
  if (!g || !sources || nsrc <= 0 || (kind != GU_QUERY_BFS && kind != GU_QUERY_BC)) return GU_ERR_INVALID;
  for (int i = 0; i < nsrc; i++)
    if (sources[i] < 0 || sources[i] >= g->n) return GU_ERR_RANGE;
  const int fam = kind == GU_QUERY_BC ? 1 : 0;
  const int k = std::min(nsrc, 256);
  const size_t n = (size_t)g->n;
  GU_HIP(hipSetDevice(g->device));
  // outputs of the trials: distance and path-count rows, or a BC vector
  int32_t *dist = nullptr;
  u64 *sig = nullptr;
  double *bc = nullptr;
  int rc = GU_OK;
  if (fam == 0) {
    if (dalloc(&dist, (size_t)k * n) != hipSuccess || dalloc(&sig, (size_t)k * n) != hipSuccess)
      rc = GU_ERR_NOMEM;
  } else if (dalloc(&bc, n) != hipSuccess) {
    rc = GU_ERR_NOMEM;
  }
  gu_graph::Tuned saved = g->tuned[fam];
  g->tuned[fam] = gu_graph::Tuned();
  int trials = 0;
  auto trial = [&](const gu_config &c, double *ms_out) -> int {
    double best = 1e300;
    for (int rep = 0; rep < 2; rep++) {
      gu_stats st{};
      int r = fam == 0 ? gu_multi_search(g, sources, k, dist, (uint64_t *)sig, &c, &st)
                       : gu_betweenness(g, sources, k, nullptr, bc, &c, &st);
      if (r) return r;
      best = std::min(best, st.device_ms);
    }
    trials++;
    *ms_out = best;
    return GU_OK;
  };
  gu_config base{};
  base.variant = GU_VARIANT_AUTO;
  double t_default = 0, t_best = 1e300, t = 0;
  gu_config best = base;
  if (rc == GU_OK) rc = trial(base, &t_default);  // warms the workspace too
  t_best = t_default;
  best.engine = resolve_engine(g, base, fam == 1 ? JOB_BC : JOB_BFS);
  // engine: per-source is a candidate on graphs without large hubs
  if (rc == GU_OK && g->max_deg <= 1024) {
    gu_config c = base;
    c.engine = best.engine == GU_ENGINE_PER_SOURCE ? GU_ENGINE_BATCHED : GU_ENGINE_PER_SOURCE;
    if ((rc = trial(c, &t)) == GU_OK && t < t_best) {
      t_best = t;
      best = c;
    }
  }
  if (rc == GU_OK && best.engine == GU_ENGINE_BATCHED) {
    best.variant = resolve_variant(g, base);
    // top-down expander
    const int vars[] = {V_SHUFFLE, V_WARP, V_CTA_WARP, V_SCAN, V_EDGE_PARALLEL};
    for (int v : vars) {
      gu_config c = best;
      c.variant = v;
      if ((rc = trial(c, &t)) != GU_OK) break;
      if (t < t_best) {
        t_best = t;
        best = c;
      }
    }
    // sources per launch, workgroup size, direction threshold
    const int widths[] = {64, 128, 256}, wgs[] = {256, 512, 1024}, alphas[] = {4, 16, 64, 256, 1024};
    const int betas[] = {16, 64, 256};
    for (int w : widths) {
      if (rc || w > ((k + 63) / 64) * 64) break;
      gu_config c = best;
      c.batch_width = w;
      if ((rc = trial(c, &t)) == GU_OK && t < t_best) {
        t_best = t;
        best = c;
      }
    }
    for (int w : wgs) {
      if (rc) break;
      gu_config c = best;
      c.wg_size = w;
      if ((rc = trial(c, &t)) == GU_OK && t < t_best) {
        t_best = t;
        best = c;
      }
    }
    for (int a : alphas) {
      if (rc) break;
      gu_config c = best;
      c.alpha = a;
      if ((rc = trial(c, &t)) == GU_OK && t < t_best) {
        t_best = t;
        best = c;
      }
    }
    for (int b : betas) {
      if (rc) break;
      gu_config c = best;
      c.beta = b;
      if ((rc = trial(c, &t)) == GU_OK && t < t_best) {
        t_best = t;
        best = c;
      }
    }
  }
  if (dist) (void)hipFree(dist);
  if (sig) (void)hipFree(sig);
  if (bc) (void)hipFree(bc);
  if (rc != GU_OK) {
    g->tuned[fam] = saved;
    return rc;
  }
  gu_graph::Tuned &tn = g->tuned[fam];
  tn.engine = best.engine;
  tn.variant = best.engine == GU_ENGINE_BATCHED ? best.variant : -1;
  tn.width = best.batch_width;
  tn.wg = best.wg_size;
  tn.alpha = best.alpha;
  tn.beta = best.beta;
  tn.sources = k;
  if (out) {
    out->engine = tn.engine;
    out->variant = tn.variant >= 0 ? tn.variant : GU_VARIANT_AUTO;
    out->batch_width = tn.width;
    out->wg_size = tn.wg;
    out->alpha = tn.alpha;
    out->beta = tn.beta;
    out->trials = trials;
    out->ms_default = t_default;
    out->ms_tuned = t_best;
    out->sources = k;
  }
  return GU_OK;
}

int gu_lbs(const int32_t *counts, int64_t S, int32_t *scan, int32_t *result, int64_t result_cap,
           int64_t *W_out, const gu_config *cfgp, gu_stats *stats) {
  if (S < 0 || (S > 0 && (!counts || !scan)) || S >= INT32_MAX) return GU_ERR_INVALID;
  gu_config cfg{};
  if (cfgp) cfg = *cfgp;
  hipStream_t st = (hipStream_t)cfg.stream;
  if (S == 0) {
    if (W_out) *W_out = 0;
    return GU_OK;
  }
  GuTraceRange tr("gu:lbs");
  int dev = 0, cus = 0;
  GU_HIP(hipGetDevice(&dev));
  GU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long long nt = lbs_scan_tiles(S);
  struct Tile {
    long long *p = nullptr;
    ~Tile() {
      if (p) (void)hipFree(p);
    }
  } tsum;
  // tile starts of the large-S expansion: W < INT_MAX items in kLbsXT tiles
  const long long ntl = (result && S > kLbsSmall) ? std::min<long long>(result_cap, INT_MAX) / kLbsXT + 2 : 0;
  GU_HIP(hipMalloc((void **)&tsum.p, sizeof(long long) * (size_t)(nt + 2) + sizeof(int) * (size_t)ntl));
  long long *total_d = tsum.p + nt + 1;
  int *tlo = reinterpret_cast<int *>(tsum.p + nt + 2);
  EventPair ev;
  GU_HIP(hipEventCreate(&ev.a));
  GU_HIP(hipEventCreate(&ev.b));
  GU_HIP(hipEventRecord(ev.a, st));
  GU_HIP(lbs_launch(counts, S, scan, result, result_cap, cus, tsum.p, total_d, tlo, st));
  GU_HIP(hipEventRecord(ev.b, st));
  long long W = 0;
  GU_HIP(hipMemcpyAsync(&W, total_d, sizeof(long long), hipMemcpyDeviceToHost, st));
  GU_HIP(hipStreamSynchronize(st));
  int rc = GU_OK;
  GU_HIP(hipEventSynchronize(ev.b));
  float ms = 0;
  GU_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
  // result == NULL asks for the scan and W only
  if (W >= INT_MAX || (result && W > result_cap)) rc = GU_ERR_CAPACITY;
  if (W_out) *W_out = W;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->device_ms = ms;
  }
  return rc;
}

int gu_lbs_noalloc(const int32_t *counts, int64_t S, int32_t *scan, int32_t *result, int64_t result_cap,
                   int64_t *W_out, const gu_config *cfgp, gu_stats *stats) {
  if (S < 0 || (S > 0 && (!counts || !scan)) || result_cap < 0) return GU_ERR_INVALID;
  if (S > kNaMaxS) return GU_ERR_CAPACITY;  // more segments than the scratch holds
  gu_config cfg{};
  if (cfgp) cfg = *cfgp;
  hipStream_t st = (hipStream_t)cfg.stream;
  if (stats) std::memset(stats, 0, sizeof(*stats));  // no device time: no events are made
  if (W_out) *W_out = 0;
  if (S == 0) return GU_OK;
  int dev = 0, cus = 0;
  GU_HIP(hipGetDevice(&dev));
  GU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  void *ptr = nullptr;
  const int rc0 = scratch_address(dev, &ptr);
  if (rc0) return rc0;
  // every scratch word a launch reads is written by the same call first, so
  // nothing needs clearing between calls
  long long *tsum = static_cast<long long *>(ptr);
  long long *total_d = tsum + lbs_scan_tiles(S) + 1;
  int *tlo = reinterpret_cast<int *>(static_cast<char *>(ptr) + kNaTsumBytes);
  // the expansion writes at most kNaMaxW items (its tile-start table)
  const long long cap = std::min<long long>(result_cap, kNaMaxW);
  GU_HIP(lbs_launch(counts, S, scan, result, cap, cus, tsum, total_d, tlo, st));
  long long W = 0;
  GU_HIP(hipMemcpyAsync(&W, total_d, sizeof(long long), hipMemcpyDeviceToHost, st));
  GU_HIP(hipStreamSynchronize(st));
  if (W_out) *W_out = W;
  if (W >= INT_MAX || (result && (W > result_cap || W > kNaMaxW))) return GU_ERR_CAPACITY;
  return GU_OK;
}

}  // extern "C"
