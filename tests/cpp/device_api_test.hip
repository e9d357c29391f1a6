// device_api_test.hip -- kernels written against the public device header
// (include/graph_utils_amd_device.hpp) only, as a caller of the device API
// would write them, with C entry points for tests/test_device_api.py.
//
// Each entry point copies its inputs to the device, launches, copies the
// outputs back and frees everything; it returns 0 or a HIP error code.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <vector>

#include "../../include/graph_utils_amd_device.hpp"

using u64 = unsigned long long;

namespace {

#define DT_HIP(call)                          \
  do {                                        \
    hipError_t e_ = (call);                   \
    if (e_ != hipSuccess) return (int)e_;     \
  } while (0)

// Device buffers freed with the scope.
struct Bufs {
  std::vector<void *> p;
  template <class T>
  hipError_t alloc(T **x, size_t count) {
    *x = nullptr;
    hipError_t e = hipMalloc((void **)x, (count ? count : 1) * sizeof(T));
    if (e == hipSuccess) p.push_back(*x);
    return e;
  }
  ~Bufs() {
    for (void *x : p) (void)hipFree(x);
  }
};

// One wave per mask (four waves per workgroup): lane l's predicate is bit l.
__global__ void k_race(const u64 *masks, int nmasks, int *out, int *lanes) {
  const int b = blockIdx.x * (blockDim.x / gu::kWave) + gu::wave_id();
  if (b >= nmasks) return;
  const int l = gu::lane_id();
  out[b * 64 + l] = gu::race_and_resolve_wave((int)((masks[b] >> l) & 1ull));
  lanes[b * 64 + l] = l + 64 * gu::wave_id();
}

// atomicAggInc: mode 0 calls it inside a branch (active lanes), mode 1 with a
// predicate, mode 2 inside a branch on a counter in LDS (then each
// workgroup's slots start at 0; the counter total goes to totals[block]).
__global__ void k_agg(const u64 *masks, int nmasks, int mode, int *ctr, unsigned *uctr, int *slots, int *totals) {
  __shared__ int s_ctr;
  if (threadIdx.x == 0) s_ctr = 0;
  __syncthreads();
  const int b = blockIdx.x * (blockDim.x / gu::kWave) + gu::wave_id();
  const int l = gu::lane_id();
  const bool p = b < nmasks && ((masks[b] >> l) & 1ull);
  int slot = -1;
  if (mode == 1) {
    slot = gu::atomic_agg_inc(uctr, p);
  } else if (p) {
    slot = gu::atomic_agg_inc(mode == 2 ? &s_ctr : ctr);
  }
  if (b < nmasks) slots[b * 64 + l] = slot;
  __syncthreads();
  if (threadIdx.x == 0) totals[blockIdx.x] = s_ctr;
}

__global__ void k_bits(const int *vals, const int *pos, const unsigned *ins, int n, int *got, unsigned *set) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  got[i] = gu::get_bit(vals[i], pos[i]);
  set[i] = gu::set_bit((unsigned)vals[i], ins[i], pos[i]);
}

// Several independent wave searches in one workgroup: wave j runs on slice j
// of the counts (offsets off[j], lengths len[j]), writing its own scan,
// result and total.
__global__ void k_lbs_waves(const int *counts, const int *off, const int *len, const long long *roff, int k,
                            int *scan, int *result, int *totals) {
  const int j = gu::wave_id();
  if (j >= k) return;
  gu::load_balance_search_wave(len[j], totals + j, counts + off[j], scan + off[j], result + roff[j]);
}

// ---- multi_search hook sets ---------------------------------------------

struct MsArgs {
  const int *R, *C;
  int n;
  int *d, *Q, *Q2, *Qs, *Q2s;
  u64 *sigma;
  float *delta;
  int *S, *E;  // Brandes stack and level starts
  int *maxd;
  gu::pitch p;
  int start, end;
};

// BFS distances: a compare-and-swap claims each vertex once, so every queue
// entry is unique (the queues hold n entries).
__global__ void k_ms_bfs(MsArgs a) {
  auto none1 = [](int) {};
  auto none2 = [](int, int) {};
  auto none3 = [](int *, int) {};
  auto none5 = [](int *, int, int, int) {};
  auto visit = [](int *d_row, int v, int w, int *Q2_row, int *Q2_len, int *, int *) {
    if (atomicCAS(&d_row[w], INT_MAX, d_row[v] + 1) == INT_MAX) gu::ms_push(Q2_row, Q2_len, w);
  };
  gu::multi_search(a.R, a.C, a.n, a.d, a.Q, a.Q2, a.Qs, a.Q2s, a.p, a.start, a.end, none1, none2, visit, none1,
                   none3, none1, none5);
}

// Shortest-path counts: claim, then every edge v->w with d[w] = d[v]+1 adds
// sigma[v] to sigma[w].  Vertices of out-degree below 8 are queued on the
// small queue (one thread each), the others on the wave queue.
__global__ void k_ms_sigma(MsArgs a) {
  u64 *s_row = gu::get_row(a.sigma, a.p.sigma);
  const int *R = a.R;
  auto none1 = [](int) {};
  auto init = [s_row](int k, int i) { s_row[k] = k == i ? 1ull : 0ull; };
  auto visit = [s_row, R](int *d_row, int v, int w, int *Q2_row, int *Q2_len, int *Q2s_row, int *Q2s_len) {
    const int dv = d_row[v];
    const int old = atomicCAS(&d_row[w], INT_MAX, dv + 1);  // w's distance, read atomically
    if (old == INT_MAX) {
      if (gu::outdegree(w, R) < 8)
        Q2s_row[atomicAdd(Q2s_len, 1)] = w;
      else
        Q2_row[atomicAdd(Q2_len, 1)] = w;
    }
    if (old == INT_MAX || old == dv + 1) atomicAdd(&s_row[w], s_row[v]);
  };
  auto none3 = [](int *, int) {};
  auto none5 = [](int *, int, int, int) {};
  gu::multi_search(a.R, a.C, a.n, a.d, a.Q, a.Q2, a.Qs, a.Q2s, a.p, a.start, a.end, none1, init, visit, none1,
                   none3, none1, none5);
}

// Diameter sampling: getMax folds every finite distance into *maxd.
__global__ void k_ms_diameter(MsArgs a) {
  int *mx = a.maxd;
  auto getmax = [mx](int dist) {
    if (dist != INT_MAX) atomicMax(mx, dist);
  };
  auto none1 = [](int) {};
  auto none2 = [](int, int) {};
  auto none3 = [](int *, int) {};
  auto none5 = [](int *, int, int, int) {};
  auto visit = [](int *d_row, int v, int w, int *Q2_row, int *Q2_len, int *, int *) {
    if (atomicCAS(&d_row[w], INT_MAX, d_row[v] + 1) == INT_MAX) gu::ms_push(Q2_row, Q2_len, w);
  };
  gu::multi_search(a.R, a.C, a.n, a.d, a.Q, a.Q2, a.Qs, a.Q2s, a.p, a.start, a.end, getmax, none2, visit, none1,
                   none3, none1, none5);
}

// Brandes: the forward keeps the BFS order as a stack S with level starts E
// (insertStack / updateEndpoints); dependencyAccum then pulls, level by level
// from the deepest, delta[v] = sum over successors w of
// sigma[v] / sigma[w] * (1 + delta[w]) (bc.cu:43-55), one thread per vertex.
__global__ void k_ms_bc(MsArgs a) {
  __shared__ int s_slen, s_elen;
  u64 *s_row = gu::get_row(a.sigma, a.p.sigma);
  float *x_row = gu::get_row(a.delta, a.p.delta);
  int *S = gu::get_row(a.S, a.p.S);
  int *E = gu::get_row(a.E, a.p.endpoints);
  const int *R = a.R, *C = a.C;
  int *sl = &s_slen, *el = &s_elen;
  auto none1 = [](int) {};
  auto init = [s_row, x_row](int k, int i) {
    s_row[k] = k == i ? 1ull : 0ull;
    x_row[k] = 0.0f;
  };
  auto visit = [s_row](int *d_row, int v, int w, int *Q2_row, int *Q2_len, int *, int *) {
    const int dv = d_row[v];
    const int old = atomicCAS(&d_row[w], INT_MAX, dv + 1);
    if (old == INT_MAX) gu::ms_push(Q2_row, Q2_len, w);
    if (old == INT_MAX || old == dv + 1) atomicAdd(&s_row[w], s_row[v]);
  };
  auto init_stack = [S, E, sl, el](int i) {
    S[0] = i;
    E[0] = 0;
    E[1] = 1;
    *sl = 1;
    *el = 2;
  };
  auto insert_stack = [S, sl](int *Q2_row, int kk) { S[*sl + kk] = Q2_row[kk]; };
  auto update_endpoints = [E, sl, el](int q2) {
    E[*el] = E[*el - 1] + q2;
    (*el)++;
    *sl += q2;
  };
  auto accum = [S, E, el, s_row, x_row, R, C](int *d_row, int i, int tid, int) {
    const int levels = *el - 1;  // level l occupies S[E[l], E[l+1])
    for (int l = levels - 2; l >= 1; l--) {
      for (int k = E[l] + tid; k < E[l + 1]; k += (int)blockDim.x) {
        const int v = S[k];
        const float sv = (float)s_row[v];
        float sum = 0.0f;
        for (int r = R[v]; r < R[v + 1]; r++) {
          const int w = C[r];
          if (d_row[w] == l + 1) sum += sv / (float)s_row[w] * (1.0f + x_row[w]);
        }
        x_row[v] = sum;
      }
      __syncthreads();
    }
    (void)i;
  };
  gu::multi_search(a.R, a.C, a.n, a.d, a.Q, a.Q2, a.Qs, a.Q2s, a.p, a.start, a.end, none1, init, visit, init_stack,
                   insert_stack, update_endpoints, accum);
}

}  // namespace

extern "C" {

int dt_race(const uint64_t *masks, int nmasks, int32_t *out, int32_t *lanes) {
  Bufs b;
  u64 *m;
  int *o, *l;
  DT_HIP(b.alloc(&m, nmasks));
  DT_HIP(b.alloc(&o, (size_t)nmasks * 64));
  DT_HIP(b.alloc(&l, (size_t)nmasks * 64));
  DT_HIP(hipMemcpy(m, masks, 8 * (size_t)nmasks, hipMemcpyHostToDevice));
  k_race<<<(nmasks + 3) / 4, 256>>>(m, nmasks, o, l);
  DT_HIP(hipGetLastError());
  DT_HIP(hipMemcpy(out, o, 4 * (size_t)nmasks * 64, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(lanes, l, 4 * (size_t)nmasks * 64, hipMemcpyDeviceToHost));
  return 0;
}

int dt_agg_inc(const uint64_t *masks, int nmasks, int mode, int32_t *slots, int32_t *total, int32_t *block_totals) {
  Bufs b;
  u64 *m;
  int *s, *c, *bt;
  unsigned *uc;
  const int blocks = (nmasks + 3) / 4;
  DT_HIP(b.alloc(&m, nmasks));
  DT_HIP(b.alloc(&s, (size_t)nmasks * 64));
  DT_HIP(b.alloc(&c, 1));
  DT_HIP(b.alloc(&uc, 1));
  DT_HIP(b.alloc(&bt, blocks));
  DT_HIP(hipMemcpy(m, masks, 8 * (size_t)nmasks, hipMemcpyHostToDevice));
  DT_HIP(hipMemset(c, 0, 4));
  DT_HIP(hipMemset(uc, 0, 4));
  k_agg<<<blocks, 256>>>(m, nmasks, mode, c, uc, s, bt);
  DT_HIP(hipGetLastError());
  DT_HIP(hipMemcpy(slots, s, 4 * (size_t)nmasks * 64, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(total, mode == 1 ? (void *)uc : (void *)c, 4, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(block_totals, bt, 4 * (size_t)blocks, hipMemcpyDeviceToHost));
  return 0;
}

int dt_bits(const int32_t *vals, const int32_t *pos, const uint32_t *ins, int n, int32_t *got, uint32_t *set) {
  Bufs b;
  int *v, *p, *g;
  unsigned *in, *s;
  DT_HIP(b.alloc(&v, n));
  DT_HIP(b.alloc(&p, n));
  DT_HIP(b.alloc(&g, n));
  DT_HIP(b.alloc(&in, n));
  DT_HIP(b.alloc(&s, n));
  DT_HIP(hipMemcpy(v, vals, 4 * (size_t)n, hipMemcpyHostToDevice));
  DT_HIP(hipMemcpy(p, pos, 4 * (size_t)n, hipMemcpyHostToDevice));
  DT_HIP(hipMemcpy(in, ins, 4 * (size_t)n, hipMemcpyHostToDevice));
  k_bits<<<(n + 255) / 256, 256>>>(v, p, in, n, g, s);
  DT_HIP(hipGetLastError());
  DT_HIP(hipMemcpy(got, g, 4 * (size_t)n, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(set, s, 4 * (size_t)n, hipMemcpyDeviceToHost));
  return 0;
}

// extract_edges_block<<<1, block>>> (form 1) or extract_edges_wave<<<1, 64>>>
// (form 0) over counts[0..S); result needs W entries (W_cap given).
int dt_lbs(const int32_t *counts, int S, int form, int block, int32_t *scan, int32_t *result, int64_t W_cap,
           int32_t *W) {
  if (S < 1 || (form == 1 && (block % 64 != 0 || block < 64 || block > 1024))) return -1;
  Bufs b;
  int *c, *s, *r, *w;
  DT_HIP(b.alloc(&c, S));
  DT_HIP(b.alloc(&s, S));
  DT_HIP(b.alloc(&r, W_cap));
  DT_HIP(b.alloc(&w, 1));
  DT_HIP(hipMemcpy(c, counts, 4 * (size_t)S, hipMemcpyHostToDevice));
  if (form == 1)
    gu::extract_edges_block<<<1, block>>>(S, c, s, r, w);
  else
    gu::extract_edges_wave<<<1, 64>>>(S, c, s, r, w);
  DT_HIP(hipGetLastError());
  DT_HIP(hipMemcpy(W, w, 4, hipMemcpyDeviceToHost));
  if (*W < 0 || *W > W_cap) return -2;
  DT_HIP(hipMemcpy(scan, s, 4 * (size_t)S, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(result, r, 4 * (size_t)*W, hipMemcpyDeviceToHost));
  return 0;
}

// k independent wave searches in one workgroup of 64*k threads (k <= 16):
// slice j = counts[off[j], off[j]+len[j]) with its result at roff[j] (the
// caller sizes result for every slice's total).
int dt_lbs_waves(const int32_t *counts, int64_t S_all, const int32_t *off, const int32_t *len,
                 const int64_t *roff, int k, int64_t W_all, int32_t *scan, int32_t *result, int32_t *totals) {
  if (k < 1 || k > 16) return -1;
  Bufs b;
  int *c, *o, *l, *s, *r, *t;
  long long *ro;
  DT_HIP(b.alloc(&c, S_all));
  DT_HIP(b.alloc(&o, k));
  DT_HIP(b.alloc(&l, k));
  DT_HIP(b.alloc(&ro, k));
  DT_HIP(b.alloc(&s, S_all));
  DT_HIP(b.alloc(&r, W_all));
  DT_HIP(b.alloc(&t, k));
  DT_HIP(hipMemcpy(c, counts, 4 * (size_t)S_all, hipMemcpyHostToDevice));
  DT_HIP(hipMemcpy(o, off, 4 * (size_t)k, hipMemcpyHostToDevice));
  DT_HIP(hipMemcpy(l, len, 4 * (size_t)k, hipMemcpyHostToDevice));
  DT_HIP(hipMemcpy(ro, roff, 8 * (size_t)k, hipMemcpyHostToDevice));
  k_lbs_waves<<<1, 64 * k>>>(c, o, l, ro, k, s, r, t);
  DT_HIP(hipGetLastError());
  DT_HIP(hipMemcpy(scan, s, 4 * (size_t)S_all, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(result, r, 4 * (size_t)W_all, hipMemcpyDeviceToHost));
  DT_HIP(hipMemcpy(totals, t, 4 * (size_t)k, hipMemcpyDeviceToHost));
  return 0;
}

// multi_search with one of the hook sets: kind 0 distances, 1 path counts,
// 2 diameter, 3 Brandes.  grid workgroups of `block` threads over the
// sources [start, end); each output row b (of grid rows) holds what
// workgroup b left there (the last source it ran).  Rows are n entries.
int dt_multi_search(const int32_t *R, const int32_t *C, int n, int64_t m, int start, int end, int kind, int grid,
                    int block, int32_t *d_rows, uint64_t *sigma_rows, float *delta_rows, int32_t *diameter) {
  if (n < 1 || grid < 1 || block < 64 || block > 1024 || block % 64) return -1;
  Bufs b;
  MsArgs a{};
  int *Rd, *Cd;
  DT_HIP(b.alloc(&Rd, (size_t)n + 1));
  DT_HIP(b.alloc(&Cd, (size_t)m));
  DT_HIP(hipMemcpy(Rd, R, 4 * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (m) DT_HIP(hipMemcpy(Cd, C, 4 * (size_t)m, hipMemcpyHostToDevice));
  const size_t rows = (size_t)grid, nn = (size_t)n;
  a.R = Rd;
  a.C = Cd;
  a.n = n;
  a.start = start;
  a.end = end;
  DT_HIP(b.alloc(&a.d, rows * nn));
  DT_HIP(b.alloc(&a.Q, rows * nn));
  DT_HIP(b.alloc(&a.Q2, rows * nn));
  DT_HIP(b.alloc(&a.Qs, rows * nn));
  DT_HIP(b.alloc(&a.Q2s, rows * nn));
  DT_HIP(b.alloc(&a.sigma, rows * nn));
  DT_HIP(b.alloc(&a.delta, rows * nn));
  DT_HIP(b.alloc(&a.S, rows * nn));
  DT_HIP(b.alloc(&a.E, rows * (nn + 2)));
  DT_HIP(b.alloc(&a.maxd, 1));
  DT_HIP(hipMemset(a.maxd, 0, 4));
  a.p.d = a.p.Q = a.p.Q2 = a.p.Q_small = a.p.Q2_small = a.p.S = nn * 4;
  a.p.sigma = nn * 8;
  a.p.delta = nn * 4;
  a.p.endpoints = (nn + 2) * 4;
  switch (kind) {
    case 0: k_ms_bfs<<<grid, block>>>(a); break;
    case 1: k_ms_sigma<<<grid, block>>>(a); break;
    case 2: k_ms_diameter<<<grid, block>>>(a); break;
    case 3: k_ms_bc<<<grid, block>>>(a); break;
    default: return -1;
  }
  DT_HIP(hipGetLastError());
  DT_HIP(hipDeviceSynchronize());
  if (d_rows) DT_HIP(hipMemcpy(d_rows, a.d, 4 * rows * nn, hipMemcpyDeviceToHost));
  if (sigma_rows) DT_HIP(hipMemcpy(sigma_rows, a.sigma, 8 * rows * nn, hipMemcpyDeviceToHost));
  if (delta_rows) DT_HIP(hipMemcpy(delta_rows, a.delta, 4 * rows * nn, hipMemcpyDeviceToHost));
  if (diameter) DT_HIP(hipMemcpy(diameter, a.maxd, 4, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
