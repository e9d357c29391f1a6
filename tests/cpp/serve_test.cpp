// serve_test.cpp -- CPU test of the query server's scheduling logic
// (graph-utils_amd/csrc/gu_serve.hpp) over a stand-in executor that computes
// each launch with the oracle (sequential BFS / path counts / FP64 Brandes)
// and sleeps in proportion to the launch size, so launches of several slots
// overlap.  Checks:
//  - the arena allocator (first fit, coalescing, exact fits);
//  - every query's result is the same run alone (one slot, one query at a
//    time) and in a mix (several slots, queries packed together and split
//    across launches, admission throttled by a small budget): distances,
//    path counts and diameters bit-exact, BC within 1e-12 relative (only
//    the grouping of the FP64 sum may differ);
//  - results equal the oracle's per-source answers;
//  - a query larger than the whole budget is rejected, a failing launch
//    fails only the queries it carried.
// Exit status 0 when every check passes.
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../../graph-utils_amd/csrc/gu_serve.hpp"
#include "../../include/graph_utils_amd.h"
#include "../../oracle/oracle.h"

namespace {

int g_fail = 0, g_checks = 0;
#define CHECK(c, ...)                          \
  do {                                         \
    g_checks++;                                \
    if (!(c)) {                                \
      g_fail++;                                \
      std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                \
      std::printf("\n");                       \
    }                                          \
  } while (0)

struct Graph {
  int n = 0;
  std::vector<int32_t> R, C;
};

Graph take(int32_t *R, int32_t *C, int32_t *F, int32_t n, int64_t m) {
  Graph g;
  g.n = n;
  g.R.assign(R, R + n + 1);
  g.C.assign(C, C + m);
  gu_host_free(R);
  gu_host_free(C);
  gu_host_free(F);
  return g;
}

struct FakeExec final : gus::Exec {
  std::vector<char> arena, host;
  std::vector<Graph> graphs;
  int poison = -1;               // a launch holding this source fails
  std::atomic<int> concurrent{0}, max_concurrent{0};
  explicit FakeExec(size_t budget, size_t host_budget = 0) : arena(budget), host(host_budget) {}
  char *arena_base() override { return arena.data(); }
  char *host_base() override { return host.data(); }
  std::atomic<int64_t> copied{0};
  int deliver_us = 0;  // time each delivery copy takes
  int to_host(int, void *dst, const void *src, size_t bytes) override {
    std::memcpy(dst, src, bytes);
    copied += (int64_t)bytes;
    return GU_OK;
  }
  int deliver(void *dst, const void *src, size_t bytes) override {
    if (deliver_us) std::this_thread::sleep_for(std::chrono::microseconds(deliver_us));
    return to_host(0, dst, src, bytes);
  }
  void busy(int k) {
    int c = ++concurrent;
    int m = max_concurrent.load();
    while (c > m && !max_concurrent.compare_exchange_weak(m, c)) {
    }
    std::this_thread::sleep_for(std::chrono::microseconds(300 + 40 * k));
    --concurrent;
  }
  int search(int, int graph, const int32_t *src, int k, int32_t *const *drows, uint64_t *const *srows,
             int32_t *ecc, int32_t *wrap) override {
    const Graph &g = graphs[(size_t)graph];
    std::vector<int32_t> d((size_t)g.n);
    std::vector<uint64_t> s((size_t)g.n);
    for (int i = 0; i < k; i++)
      if (src[i] == poison) return GU_ERR_HIP;
    for (int i = 0; i < k; i++) {
      or_sigma(g.R.data(), g.C.data(), g.n, src[i], d.data(), s.data());
      if (drows) std::memcpy(drows[i], d.data(), sizeof(int32_t) * (size_t)g.n);
      if (srows) std::memcpy(srows[i], s.data(), sizeof(uint64_t) * (size_t)g.n);
      if (wrap) wrap[i] = srows ? or_sigma_wraps(g.R.data(), g.C.data(), g.n, src[i]) : 0;
      if (ecc) {
        int e = 0;
        for (int v = 0; v < g.n; v++)
          if (d[(size_t)v] != INT_MAX) e = std::max(e, (int)d[(size_t)v]);
        ecc[i] = e;
      }
    }
    busy(k);
    return GU_OK;
  }
  int bc(int, int graph, const int32_t *src, int k, float *const *xrows, double *const *targets,
         int32_t *wrap) override {
    const Graph &g = graphs[(size_t)graph];
    std::vector<double> delta((size_t)g.n), acc((size_t)g.n, 0.0);
    for (int i = 0; i < k; i++) {
      if (src[i] == poison) return GU_ERR_HIP;
    }
    for (int i = 0; i < k; i++) {
      or_brandes_f64(g.R.data(), g.C.data(), g.n, src[i], nullptr, nullptr, delta.data());
      if (wrap) wrap[i] = or_sigma_wraps(g.R.data(), g.C.data(), g.n, src[i]);
      if (xrows && xrows[i])
        for (int v = 0; v < g.n; v++) xrows[i][v] = (float)delta[(size_t)v];
      for (int v = 0; v < g.n; v++) acc[(size_t)v] += delta[(size_t)v];
      if (i + 1 == k || targets[i + 1] != targets[i]) {
        if (targets[i])
          for (int v = 0; v < g.n; v++) targets[i][v] += acc[(size_t)v];
        std::fill(acc.begin(), acc.end(), 0.0);
      }
    }
    busy(k);
    return GU_OK;
  }
  int zero(int, void *p, size_t bytes) override {
    std::memset(p, 0, bytes);
    return GU_OK;
  }
};

struct QSpec {
  int graph, cls, nsrc;
  bool want_delta;
  std::vector<int32_t> src;
};

struct QOut {
  int status = 0, diameter = 0, launches = 0, flags = 0;
  std::vector<int32_t> dist;
  std::vector<uint64_t> sigma;
  std::vector<float> delta;
  std::vector<double> bc;
};

QOut copy_out(const gus::Result &r, int n) {
  QOut o;
  o.status = r.status;
  o.diameter = r.diameter;
  o.launches = r.launches;
  o.flags = r.flags;
  const size_t k = (size_t)r.nsrc, N = (size_t)n;
  // host-delivered queries: results in the host arena, device pointers null
  const int32_t *d = r.h_dist ? r.h_dist : r.dist;
  const uint64_t *s = r.h_sigma ? r.h_sigma : r.sigma;
  const float *x = r.h_delta ? r.h_delta : r.delta;
  const double *b = r.h_bc ? r.h_bc : r.bc;
  if (d) o.dist.assign(d, d + k * N);
  if (s) o.sigma.assign(s, s + k * N);
  if (x) o.delta.assign(x, x + k * N);
  if (b) o.bc.assign(b, b + N);
  return o;
}

void test_arena() {
  gus::Arena a(4096);
  size_t o1, o2, o3, o4;
  CHECK(a.alloc(1000, &o1) && o1 == 0, "first block at 0");
  CHECK(a.alloc(1024, &o2) && o2 == 1024, "second block after the rounded first");
  CHECK(a.alloc(2048, &o3) && o3 == 2048, "exact fit of the rest");
  CHECK(!a.alloc(1, &o4), "full arena refuses");
  a.release(o2, 1024);
  CHECK(a.alloc(700, &o4) && o4 == 1024, "hole reused");
  a.release(o4, 700);
  a.release(o1, 1000);
  CHECK(a.free_blocks() == 1, "neighbouring holes coalesce (%zu)", a.free_blocks());
  CHECK(a.alloc(2048, &o4) && o4 == 0, "coalesced hole holds 2048");
  a.release(o4, 2048);
  a.release(o3, 2048);
  CHECK(a.used() == 0 && a.free_blocks() == 1, "empty arena is one block");
  CHECK(a.peak() == 4096, "peak %zu", a.peak());
}

std::vector<QSpec> make_queries(const std::vector<Graph> &gs, int count, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::vector<QSpec> qs;
  for (int i = 0; i < count; i++) {
    QSpec q;
    q.graph = (int)(rng() % gs.size());
    const int r = (int)(rng() % 10);
    q.cls = r < 5 ? gus::C_BFS_SIGMA : (r < 6 ? gus::C_BFS : (r < 8 ? gus::C_BC : gus::C_DIAM));
    // log-uniform sizes in [1, 40]
    q.nsrc = std::max(1, (int)std::floor(std::exp(std::log(40.0) * (double)(rng() % 10000) / 10000.0)));
    q.want_delta = q.cls == gus::C_BC && (rng() % 2);
    for (int j = 0; j < q.nsrc; j++) q.src.push_back((int32_t)(rng() % (uint64_t)gs[(size_t)q.graph].n));
    qs.push_back(q);
  }
  return qs;
}

// results delivered to the host: 0 for no query, 1 for every query, 2 for
// the queries whose first source is odd (launches mixing both kinds)
int g_to_host = 0;

gus::Spec spec_of(const QSpec &q) {
  gus::Spec s;
  s.to_host = g_to_host == 1 || (g_to_host == 2 && (q.src[0] & 1));
  s.graph = q.graph;
  s.cls = q.cls;
  s.nsrc = q.nsrc;
  s.sources = q.src.data();
  s.want_delta = q.want_delta;
  return s;
}

std::vector<QOut> run_alone(const std::vector<Graph> &gs, const std::vector<QSpec> &qs, size_t budget) {
  FakeExec ex(budget);
  ex.graphs = gs;
  gus::Server srv(&ex, budget, 1, 256);
  for (const Graph &g : gs) {
    int id;
    srv.add_graph(g.n, &id);
  }
  std::vector<QOut> out;
  for (const QSpec &q : qs) {
    int64_t t;
    CHECK(srv.submit(spec_of(q), &t) == 0, "submit alone");
    gus::Result r;
    CHECK(srv.wait(t, -1, &r) == 0, "wait alone");
    CHECK(r.launches == 1, "alone: one launch per query (%d)", r.launches);
    out.push_back(copy_out(r, gs[(size_t)q.graph].n));
    srv.release(t);
  }
  return out;
}

std::vector<QOut> run_mix(const std::vector<Graph> &gs, const std::vector<QSpec> &qs, size_t budget,
                          int slots, int max_pack, int *max_conc, gus::Stats *st, size_t host_budget = 0) {
  FakeExec ex(budget, host_budget);
  ex.graphs = gs;
  gus::Server srv(&ex, budget, slots, max_pack, host_budget);
  for (const Graph &g : gs) {
    int id;
    srv.add_graph(g.n, &id);
  }
  std::vector<QOut> out(qs.size());
  std::vector<int64_t> tickets(qs.size());
  std::mutex mu;
  std::unordered_map<int64_t, size_t> index;
  // two client threads submit interleaved halves with small gaps
  auto client = [&](int part) {
    std::mt19937 rng(77 + part);
    for (size_t i = (size_t)part; i < qs.size(); i += 2) {
      int64_t t;
      std::lock_guard<std::mutex> lk(mu);
      CHECK(srv.submit(spec_of(qs[i]), &t) == 0, "submit mix");
      index[t] = i;
    }
    (void)rng;
  };
  std::thread c0(client, 0), c1(client, 1);
  c0.join();
  c1.join();
  for (size_t done = 0; done < qs.size(); done++) {
    int64_t t;
    gus::Result r;
    CHECK(srv.next_done(20000, &t, &r) == 0, "next_done");
    size_t i;
    {
      std::lock_guard<std::mutex> lk(mu);
      i = index[t];
    }
    out[i] = copy_out(r, gs[(size_t)qs[i].graph].n);
    CHECK(srv.release(t) == 0, "release");
  }
  *max_conc = ex.max_concurrent.load();
  *st = srv.stats();
  return out;
}

void compare(const std::vector<Graph> &gs, const std::vector<QSpec> &qs, const std::vector<QOut> &a,
             const std::vector<QOut> &b, const char *label) {
  for (size_t i = 0; i < qs.size(); i++) {
    const QSpec &q = qs[i];
    CHECK(a[i].status == 0 && b[i].status == 0, "%s q%zu status %d %d", label, i, a[i].status, b[i].status);
    CHECK(a[i].dist == b[i].dist, "%s q%zu distances differ", label, i);
    CHECK(a[i].sigma == b[i].sigma, "%s q%zu path counts differ", label, i);
    CHECK(a[i].delta == b[i].delta, "%s q%zu dependency rows differ", label, i);
    CHECK(a[i].diameter == b[i].diameter, "%s q%zu diameter %d vs %d", label, i, a[i].diameter, b[i].diameter);
    CHECK(a[i].flags == b[i].flags, "%s q%zu flags %d vs %d", label, i, a[i].flags, b[i].flags);
    CHECK(a[i].bc.size() == b[i].bc.size(), "%s q%zu bc size", label, i);
    double worst = 0;
    for (size_t v = 0; v < a[i].bc.size() && v < b[i].bc.size(); v++) {
      const double x = a[i].bc[v], y = b[i].bc[v];
      if (x != y) worst = std::max(worst, std::fabs(x - y) / std::max(std::fabs(x), 1e-300));
    }
    CHECK(worst <= 1e-12, "%s q%zu bc rel diff %g", label, i, worst);
    // against the oracle directly
    const Graph &g = gs[(size_t)q.graph];
    std::vector<int32_t> d((size_t)g.n);
    std::vector<uint64_t> s((size_t)g.n);
    int diam = 0;
    for (int j = 0; j < q.nsrc; j++) {
      or_sigma(g.R.data(), g.C.data(), g.n, q.src[(size_t)j], d.data(), s.data());
      for (int v = 0; v < g.n; v++) {
        if (d[(size_t)v] != INT_MAX) diam = std::max(diam, (int)d[(size_t)v]);
        if (!b[i].dist.empty() && b[i].dist[(size_t)j * g.n + v] != d[(size_t)v]) {
          CHECK(false, "%s q%zu source %d vertex %d distance vs oracle", label, i, j, v);
          j = q.nsrc;
          break;
        }
        if (!b[i].sigma.empty() && b[i].sigma[(size_t)j * g.n + v] != s[(size_t)v]) {
          CHECK(false, "%s q%zu source %d vertex %d path count vs oracle", label, i, j, v);
          j = q.nsrc;
          break;
        }
      }
    }
    if (q.cls == gus::C_DIAM) CHECK(b[i].diameter == diam, "%s q%zu diameter vs oracle", label, i);
    if (q.cls == gus::C_BC) {
      std::vector<double> ref((size_t)g.n);
      or_bc_f64(g.R.data(), g.C.data(), g.n, q.src.data(), q.nsrc, ref.data());
      double w = 0;
      for (int v = 0; v < g.n; v++)
        if (ref[(size_t)v] > 0) w = std::max(w, std::fabs(b[i].bc[(size_t)v] - ref[(size_t)v]) / ref[(size_t)v]);
        else CHECK(b[i].bc[(size_t)v] == 0.0, "%s q%zu bc nonzero where oracle is 0", label, i);
      CHECK(w <= 1e-12, "%s q%zu bc vs oracle %g", label, i, w);
    }
  }
}

}  // namespace

int main() {
  test_arena();
  std::vector<Graph> gs;
  {
    int32_t *R, *C, *F, n;
    int64_t m;
    if (gu_gen_er(300, 900, 5, &R, &C, &F, &n, &m) != GU_OK) return 2;
    gs.push_back(take(R, C, F, n, m));
    if (gu_gen_grid(24, 0.1, 6, 1, &R, &C, &F, &n, &m) != GU_OK) return 2;
    gs.push_back(take(R, C, F, n, m));
    if (gu_gen_rmat(9, 8, 0.57, 0.19, 0.19, 7, &R, &C, &F, &n, &m) != GU_OK) return 2;
    gs.push_back(take(R, C, F, n, m));
  }
  const std::vector<QSpec> qs = make_queries(gs, 90, 2024);
  size_t largest = 0;
  for (const QSpec &q : qs)
    largest = std::max(largest, gus::query_bytes(q.cls, q.nsrc, gs[(size_t)q.graph].n, q.want_delta));
  const std::vector<QOut> alone = run_alone(gs, qs, largest * 2);
  int conc = 0;
  gus::Stats st;
  // a budget of three of the largest queries forces admission to wait
  const std::vector<QOut> mix = run_mix(gs, qs, largest * 3, 4, 16, &conc, &st);
  compare(gs, qs, alone, mix, "mix");
  CHECK(conc >= 2, "launches overlapped (max %d in flight)", conc);
  CHECK(st.completed == (int64_t)qs.size(), "completed %lld", (long long)st.completed);
  CHECK(st.hbm_peak <= (int64_t)(largest * 3), "peak %lld within budget", (long long)st.hbm_peak);
  CHECK(st.hbm_in_use == 0, "everything released");
  int64_t srcs = 0;
  for (const QSpec &q : qs) srcs += q.nsrc;
  CHECK(st.sources == srcs, "every source launched once (%lld vs %lld)", (long long)st.sources, (long long)srcs);
  std::printf("mix: %lld launches for %zu queries (%lld sources), max %d in flight, peak %lld of %zu bytes\n",
              (long long)st.launches, qs.size(), (long long)srcs, conc, (long long)st.hbm_peak, largest * 3);
  CHECK(st.launches < (int64_t)qs.size(), "queries shared launches");

  // rejection and failure isolation
  {
    FakeExec ex(1 << 20);
    ex.graphs = gs;
    ex.poison = 7;
    gus::Server srv(&ex, 1 << 20, 2, 7);  // the poisoned query fills a launch
    int id;
    srv.add_graph(gs[0].n, &id);
    std::vector<int32_t> big(2000, 1), bad = {1, 2, 3, 4, 5, 6, 7}, good = {8, 9, 10};
    gus::Spec s;
    s.graph = 0;
    s.cls = gus::C_BFS_SIGMA;
    s.nsrc = (int)big.size();
    s.sources = big.data();
    int64_t tb, tx, tg;
    CHECK(srv.submit(s, &tb) == 0, "submit big");
    s.nsrc = (int)bad.size();
    s.sources = bad.data();
    CHECK(srv.submit(s, &tx) == 0, "submit poisoned");
    s.nsrc = (int)good.size();
    s.sources = good.data();
    CHECK(srv.submit(s, &tg) == 0, "submit good");
    gus::Result r;
    CHECK(srv.wait(tb, 5000, &r) == 0 && r.status == gus::kErrCapacity, "oversized query rejected (%d)", r.status);
    CHECK(srv.wait(tx, 5000, &r) == 0 && r.status == GU_ERR_HIP, "poisoned query fails (%d)", r.status);
    CHECK(srv.wait(tg, 5000, &r) == 0 && r.status == 0, "other query unaffected (%d)", r.status);
    CHECK(srv.release(tb) == 0 && srv.release(tx) == 0 && srv.release(tg) == 0, "release all");
    int64_t tt;
    CHECK(srv.submit(gus::Spec(), &tt) == gus::kErrInvalid, "empty query refused");
    s.graph = 5;
    CHECK(srv.submit(s, &tt) == gus::kErrInvalid, "unknown graph refused");
    s.graph = 0;
    std::vector<int32_t> oob = {gs[0].n};
    s.sources = oob.data();
    s.nsrc = 1;
    CHECK(srv.submit(s, &tt) == gus::kErrInvalid, "out-of-range source refused");
  }
  // results delivered to the host: the same answers, the device block given
  // back once a query is done, admission throttled by the host budget too
  {
    g_to_host = 1;
    gus::Stats hst;
    int hconc = 0;
    const std::vector<QOut> hmix = run_mix(gs, qs, largest * 3, 4, 16, &hconc, &hst, largest * 4);
    compare(gs, qs, alone, hmix, "host mix");
    CHECK(hst.completed == (int64_t)qs.size(), "host mix completed %lld", (long long)hst.completed);
    CHECK(hst.hbm_in_use == 0 && hst.host_in_use == 0, "host mix: everything released");
    CHECK(hst.host_peak > 0 && hst.host_peak <= (int64_t)(largest * 4), "host peak %lld", (long long)hst.host_peak);
    // host and device destinations in the same launches
    g_to_host = 2;
    const std::vector<QOut> bmix = run_mix(gs, qs, largest * 3, 4, 16, &hconc, &hst, largest * 4);
    g_to_host = 0;
    compare(gs, qs, alone, bmix, "host and device mix");
    CHECK(hst.completed == (int64_t)qs.size(), "host and device mix completed %lld", (long long)hst.completed);
    CHECK(hst.hbm_in_use == 0 && hst.host_in_use == 0, "host and device mix: everything released");

    FakeExec ex(1 << 20, 1 << 20);
    ex.graphs = gs;
    gus::Server srv(&ex, 1 << 20, 2, 8, 1 << 20);
    int id;
    srv.add_graph(gs[0].n, &id);
    std::vector<int32_t> src = {1, 2, 3};
    gus::Spec s;
    s.graph = 0;
    s.cls = gus::C_BFS_SIGMA;
    s.nsrc = 3;
    s.sources = src.data();
    s.to_host = true;
    int64_t t;
    CHECK(srv.submit(s, &t) == 0, "submit host query");
    gus::Result r;
    CHECK(srv.wait(t, 5000, &r) == 0 && r.status == 0, "host query done");
    gus::Stats a = srv.stats();
    CHECK(!r.dist && !r.sigma && r.h_dist && r.h_sigma, "host query: host pointers only");
    CHECK(a.hbm_in_use == 0 && a.host_in_use > 0, "device block back at completion (%lld, %lld)",
          (long long)a.hbm_in_use, (long long)a.host_in_use);
    CHECK(ex.copied.load() == (int64_t)(3 * gs[0].n * 12), "copied %lld bytes", (long long)ex.copied.load());
    CHECK(srv.release(t) == 0 && srv.stats().host_in_use == 0, "host block released");
    // deliveries still queued when the server stops are made before stop()
    // returns: slow copies, a launch per source, and a stop while the
    // workers are ahead of the delivery thread
    {
      FakeExec slow(1 << 20, 1 << 20);
      slow.graphs = gs;
      slow.deliver_us = 4000;
      gus::Server ss(&slow, 1 << 20, 2, 1, 1 << 20);
      ss.add_graph(gs[0].n, &id);
      std::vector<int32_t> many = {1, 2, 3, 4, 5, 6};
      gus::Spec sp = s;
      sp.sources = many.data();
      sp.nsrc = 6;
      int64_t tq;
      CHECK(ss.submit(sp, &tq) == 0, "submit slow host query");
      std::this_thread::sleep_for(std::chrono::milliseconds(6));
      ss.stop();
      const gus::Stats st = ss.stats();
      CHECK(st.sources >= 1 && slow.copied.load() == st.sources * (int64_t)gs[0].n * 12,
            "every launched source delivered by stop (%lld sources, %lld bytes)", (long long)st.sources,
            (long long)slow.copied.load());
    }
    // a server without a host arena cannot hold host results
    gus::Server none(&ex, 1 << 20, 1, 8);
    none.add_graph(gs[0].n, &id);
    CHECK(none.submit(s, &t) == 0 && none.wait(t, 5000, &r) == 0 && r.status == gus::kErrCapacity,
          "host query without a host arena is refused (%d)", r.status);
    none.release(t);
  }

  // wrapped path counts are flagged per query: on the intact 40 x 40 grid
  // the counts from a corner reach C(78, 39) > 2^64, those from the centre
  // stay below 2^38
  {
    int32_t *R, *C, *F, n;
    int64_t m;
    if (gu_gen_grid(40, 0.0, 1, 0, &R, &C, &F, &n, &m) != GU_OK) return 2;
    std::vector<Graph> wg = {take(R, C, F, n, m)};
    FakeExec ex(64 << 20);
    ex.graphs = wg;
    gus::Server srv(&ex, 64 << 20, 2, 2);  // packs of 2: the queries share launches
    int id;
    srv.add_graph(wg[0].n, &id);
    std::vector<int32_t> corner = {0}, centre = {20 * 40 + 20}, both = {20 * 40 + 20, 39};
    const int cls[] = {gus::C_BFS_SIGMA, gus::C_BC, gus::C_BFS, gus::C_DIAM};
    for (int c : cls) {
      int64_t t[3];
      const std::vector<int32_t> *in[3] = {&corner, &centre, &both};
      for (int j = 0; j < 3; j++) {
        gus::Spec s;
        s.graph = 0;
        s.cls = c;
        s.nsrc = (int)in[j]->size();
        s.sources = in[j]->data();
        CHECK(srv.submit(s, &t[j]) == 0, "submit wrap query");
      }
      const bool counts = c == gus::C_BFS_SIGMA || c == gus::C_BC;
      const int want[3] = {counts ? gus::kFlagWrapped : 0, 0, counts ? gus::kFlagWrapped : 0};
      for (int j = 0; j < 3; j++) {
        gus::Result r;
        CHECK(srv.wait(t[j], 5000, &r) == 0 && r.status == 0, "wrap query done");
        CHECK(r.flags == want[j], "class %d query %d: flags %d, want %d", c, j, r.flags, want[j]);
        CHECK(srv.release(t[j]) == 0, "release wrap query");
      }
    }
  }

  // a client waiting on a ticket that another client releases: the waiter
  // looks the ticket up again and gets kErrInvalid (no access to the freed
  // query)
  for (int rep = 0; rep < 50; rep++) {
    FakeExec ex(1 << 20);
    ex.graphs = gs;
    gus::Server srv(&ex, 1 << 20, 1, 8);
    int id;
    srv.add_graph(gs[0].n, &id);
    std::vector<int32_t> src = {1, 2, 3};
    gus::Spec s;
    s.graph = 0;
    s.cls = gus::C_BFS_SIGMA;
    s.nsrc = 3;
    s.sources = src.data();
    int64_t t;
    CHECK(srv.submit(s, &t) == 0, "submit for the release race");
    std::atomic<int> got{1};
    std::thread waiter([&] {
      gus::Result r;
      got = srv.wait(t, 5000, &r);
    });
    int done = 0;
    while (!done) {
      srv.poll(t, &done);
      std::this_thread::yield();
    }
    const int rr = srv.release(t);
    waiter.join();
    CHECK(rr == 0, "release while another thread waits");
    CHECK(got == 0 || got == gus::kErrInvalid, "waiter returned %d", got.load());
  }
  std::printf("serve_test: %d checks, %d failed\n", g_checks, g_fail);
  return g_fail ? 1 : 0;
}
