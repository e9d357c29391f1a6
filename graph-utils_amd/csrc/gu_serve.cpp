// gu_serve.cpp -- C ABI of the query server: the scheduler of gu_serve.hpp
// over the HIP engine.  Each slot owns a HIP stream and, per registered
// graph, a shared handle (gu_graph_share) with its own search workspace, so
// launches of different slots overlap on the device while reading one
// resident copy of the graph.  A server with a host arena also has one
// delivery stream, on which every device-to-host copy of results runs.  The
// output arenas are allocated when the server is created and each slot's
// search workspace when a graph is registered; nothing is allocated per
// query.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "../../include/graph_utils_amd.h"
#include "gu_internal.hpp"
#include "gu_serve.hpp"

namespace {

struct GpuExec final : gus::Exec {
  int device = 0;
  char *arena = nullptr;
  char *host = nullptr;  // pinned host arena (results delivered to the host)
  int slots = 0;
  gu_config base{};
  std::vector<hipStream_t> streams;
  hipStream_t dstream = nullptr;  // the delivery thread's copies to the host
  gu_graph_t views[gus::kMaxGraphs][64] = {};
  int ngraphs = 0;

  char *arena_base() override { return arena; }
  void thread_init(int) override { (void)hipSetDevice(device); }
  gu_config cfg(int slot) const {
    gu_config c = base;
    c.stream = (void *)streams[(size_t)slot];
    return c;
  }
  int search(int slot, int graph, const int32_t *src, int k, int32_t *const *drows,
             uint64_t *const *srows, int32_t *ecc, int32_t *wrap) override {
    gu_config c = cfg(slot);
    return gu_search_rows_wrap(views[graph][slot], src, k, drows, srows, ecc, wrap, &c, nullptr);
  }
  int bc(int slot, int graph, const int32_t *src, int k, float *const *xrows,
         double *const *targets, int32_t *wrap) override {
    gu_config c = cfg(slot);
    return gu_bc_rows_wrap(views[graph][slot], src, k, xrows, targets, wrap, &c, nullptr);
  }
  int zero(int slot, void *p, size_t bytes) override {
    return hipMemsetAsync(p, 0, bytes, streams[(size_t)slot]) == hipSuccess ? GU_OK : GU_ERR_HIP;
  }
  char *host_base() override { return host; }
  int to_host(int slot, void *dst, const void *src, size_t bytes) override {
    hipStream_t st = streams[(size_t)slot];
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return GU_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? GU_OK : GU_ERR_HIP;
  }
  int deliver(void *dst, const void *src, size_t bytes) override {
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, dstream) != hipSuccess) return GU_ERR_HIP;
    return hipStreamSynchronize(dstream) == hipSuccess ? GU_OK : GU_ERR_HIP;
  }
  void free_all() {
    for (hipStream_t x : streams) (void)hipStreamDestroy(x);
    streams.clear();
    if (dstream) (void)hipStreamDestroy(dstream);
    dstream = nullptr;
    if (arena) (void)hipFree(arena);
    if (host) (void)hipHostFree(host);
    arena = nullptr;
    host = nullptr;
  }
};

}  // namespace

struct gu_server {
  GpuExec ex;
  gus::Server *srv = nullptr;
};

extern "C" {

int gu_server_create(const gu_server_config *cfgp, gu_server_t *out) {
  if (!out) return GU_ERR_INVALID;
  *out = nullptr;
  gu_server_config c{};
  if (cfgp) c = *cfgp;
  if (c.slots < 0 || c.slots > 64 || c.max_pack < 0 || c.max_pack > 256 || c.hbm_budget < 0 ||
      c.host_budget < 0)
    return GU_ERR_INVALID;
  const int slots = c.slots ? c.slots : 4;
  const int max_pack = c.max_pack ? c.max_pack : 256;
  gu_server *s = new (std::nothrow) gu_server;
  if (!s) return GU_ERR_NOMEM;
  if (hipGetDevice(&s->ex.device) != hipSuccess) {
    delete s;
    return GU_ERR_NODEVICE;
  }
  size_t budget = (size_t)c.hbm_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return GU_ERR_HIP;
    }
    budget = fr / 4;
  }
  if (hipMalloc((void **)&s->ex.arena, budget) != hipSuccess) {
    delete s;
    return GU_ERR_NOMEM;
  }
  if (c.host_budget > 0 &&
      hipHostMalloc((void **)&s->ex.host, (size_t)c.host_budget, hipHostMallocDefault) != hipSuccess) {
    s->ex.free_all();
    delete s;
    return GU_ERR_NOMEM;
  }
  if (c.host_budget > 0 && hipStreamCreateWithFlags(&s->ex.dstream, hipStreamNonBlocking) != hipSuccess) {
    s->ex.dstream = nullptr;
    s->ex.free_all();
    delete s;
    return GU_ERR_HIP;
  }
  s->ex.slots = slots;
  s->ex.base.variant = c.variant ? c.variant : GU_VARIANT_AUTO;
  s->ex.base.engine = c.engine;
  for (int i = 0; i < slots; i++) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      s->ex.free_all();
      delete s;
      return GU_ERR_HIP;
    }
    s->ex.streams.push_back(st);
  }
  s->srv = new (std::nothrow) gus::Server(&s->ex, budget, slots, max_pack, (size_t)c.host_budget);
  if (!s->srv) {
    s->ex.free_all();
    delete s;
    return GU_ERR_NOMEM;
  }
  *out = s;
  return GU_OK;
}

int gu_server_add_graph(gu_server_t s, gu_graph_t g, int32_t *graph_id) {
  if (!s || !g || !graph_id) return GU_ERR_INVALID;
  if (s->ex.ngraphs >= gus::kMaxGraphs) return GU_ERR_CAPACITY;
  int32_t n = 0;
  int rc = gu_graph_arrays(g, nullptr, nullptr, nullptr, &n, nullptr);
  if (rc) return rc;
  const int gid = s->ex.ngraphs;
  for (int i = 0; i < s->ex.slots; i++) {
    rc = gu_graph_share(g, &s->ex.views[gid][i]);
    if (!rc) {
      rc = gu_reserve_launches(s->ex.views[gid][i], &s->ex.base);
      if (rc) gu_graph_destroy(s->ex.views[gid][i]);
    }
    if (rc) {
      for (int j = 0; j < i; j++) gu_graph_destroy(s->ex.views[gid][j]);
      return rc;
    }
  }
  s->ex.ngraphs++;
  int id = -1;
  rc = s->srv->add_graph(n, &id);
  if (rc) return GU_ERR_INVALID;
  *graph_id = id;
  return GU_OK;
}

int gu_server_submit(gu_server_t s, const gu_query *q, int64_t *ticket) {
  if (!s || !q || !ticket) return GU_ERR_INVALID;
  gus::Spec sp;
  sp.graph = q->graph;
  sp.nsrc = q->nsrc;
  sp.sources = q->sources;
  sp.to_host = (q->flags & GU_QF_HOST) != 0;
  if (sp.to_host && !s->ex.host) return GU_ERR_INVALID;  // the server has no host arena
  switch (q->kind) {
    case GU_QUERY_BFS: sp.cls = (q->flags & GU_QF_SIGMA) ? gus::C_BFS_SIGMA : gus::C_BFS; break;
    case GU_QUERY_BC: sp.cls = gus::C_BC; sp.want_delta = (q->flags & GU_QF_DELTA) != 0; break;
    case GU_QUERY_DIAMETER: sp.cls = gus::C_DIAM; break;
    default: return GU_ERR_INVALID;
  }
  return s->srv->submit(sp, ticket);
}

static void to_c(const gus::Result &r, gu_query_result *o) {
  o->status = r.status;
  o->kind = r.cls == gus::C_BC ? GU_QUERY_BC : (r.cls == gus::C_DIAM ? GU_QUERY_DIAMETER : GU_QUERY_BFS);
  o->nsrc = r.nsrc;
  o->diameter = r.diameter;
  o->launches = r.launches;
  o->flags = (r.flags & gus::kFlagWrapped) ? GU_QR_SIGMA_WRAPPED : 0;
  o->dist_dev = r.dist;
  o->sigma_dev = r.sigma;
  o->delta_dev = r.delta;
  o->bc_dev = r.bc;
  o->dist_host = r.h_dist;
  o->sigma_host = r.h_sigma;
  o->delta_host = r.h_delta;
  o->bc_host = r.h_bc;
  o->submit_ms = r.submit_ms;
  o->admit_ms = r.admit_ms;
  o->first_launch_ms = r.first_launch_ms;
  o->done_ms = r.done_ms;
  o->hbm_bytes = (int64_t)r.bytes;
}

int gu_server_wait(gu_server_t s, int64_t ticket, int32_t timeout_ms, gu_query_result *res) {
  if (!s || !res) return GU_ERR_INVALID;
  gus::Result r;
  int rc = s->srv->wait(ticket, timeout_ms, &r);
  if (rc) return rc;
  to_c(r, res);
  return GU_OK;
}

int gu_server_next_done(gu_server_t s, int32_t timeout_ms, int64_t *ticket, gu_query_result *res) {
  if (!s || !ticket || !res) return GU_ERR_INVALID;
  gus::Result r;
  int rc = s->srv->next_done(timeout_ms, ticket, &r);
  if (rc) return rc;
  to_c(r, res);
  return GU_OK;
}

int gu_server_poll(gu_server_t s, int64_t ticket, int32_t *done) {
  if (!s || !done) return GU_ERR_INVALID;
  int d = 0;
  int rc = s->srv->poll(ticket, &d);
  *done = d;
  return rc;
}

int gu_server_release(gu_server_t s, int64_t ticket) {
  if (!s) return GU_ERR_INVALID;
  return s->srv->release(ticket);
}

int gu_server_now_ms(gu_server_t s, double *ms) {
  if (!s || !ms) return GU_ERR_INVALID;
  *ms = s->srv->now_ms();
  return GU_OK;
}

int gu_server_get_stats(gu_server_t s, gu_server_stats *out) {
  if (!s || !out) return GU_ERR_INVALID;
  gus::Stats st = s->srv->stats();
  out->submitted = st.submitted;
  out->completed = st.completed;
  out->rejected = st.rejected;
  out->launches = st.launches;
  out->sources = st.sources;
  out->hbm_in_use = st.hbm_in_use;
  out->hbm_peak = st.hbm_peak;
  out->hbm_budget = st.hbm_budget;
  out->waiting = st.waiting;
  out->active = st.active;
  out->host_in_use = st.host_in_use;
  out->host_peak = st.host_peak;
  out->host_budget = st.host_budget;
  return GU_OK;
}

int gu_server_destroy(gu_server_t s) {
  if (!s) return GU_ERR_INVALID;
  delete s->srv;  // stops the workers after their current launches
  (void)hipSetDevice(s->ex.device);
  for (int g = 0; g < s->ex.ngraphs; g++)
    for (int i = 0; i < s->ex.slots; i++) gu_graph_destroy(s->ex.views[g][i]);
  s->ex.free_all();
  delete s;
  return GU_OK;
}

}  // extern "C"
