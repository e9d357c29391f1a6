// ref_parse_shim.cpp -- C entry points over the reference's own graph readers.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's unmodified parse.cpp and util.cpp (read in place from the
// reference checkout) into oracle/_ref/libref_parse.so, so tests can compare
// the product's readers with parse_metis, parse_market and parse_snap
// (parse.cpp:115-434) themselves, and its reverse and DAG reachability with
// reverse() and get_reachability_dag() (graph-utils/reverse.cpp,
// dag_reachability.cpp).
// Nothing here restates the reader; it only copies host_graph's vectors out.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "parse.h"
#include "graph-utils/dag_reachability.h"
#include "graph-utils/reverse.h"

extern "C" {

// Parses `path` with the reference's parse_metis and returns malloc'd copies
// of R (n+1), C (m) and F (m).  Release them with ref_free.  The reference
// exits the process on unreadable or weighted files, so callers only pass
// files that it accepts.
int ref_parse_metis(const char *path, int32_t *n, int64_t *m, int32_t **R,
                    int32_t **C, int32_t **F, int32_t *directed) {
  std::vector<char> p(path, path + std::strlen(path) + 1);
  host_graph g = parse_metis(p.data());
  *n = g.n;
  *m = g.m;
  *directed = g.directed ? 1 : 0;
  *R = (int32_t *)std::malloc(sizeof(int32_t) * (g.R.size() ? g.R.size() : 1));
  *C = (int32_t *)std::malloc(sizeof(int32_t) * (g.C.size() ? g.C.size() : 1));
  *F = (int32_t *)std::malloc(sizeof(int32_t) * (g.F.size() ? g.F.size() : 1));
  if (!g.R.empty()) std::memcpy(*R, g.R.data(), sizeof(int32_t) * g.R.size());
  if (!g.C.empty()) std::memcpy(*C, g.C.data(), sizeof(int32_t) * g.C.size());
  if (!g.F.empty()) std::memcpy(*F, g.F.data(), sizeof(int32_t) * g.F.size());
  return 0;
}

static void copy_out(const host_graph &g, int32_t *n, int64_t *m, int32_t **R, int32_t **C, int32_t **F) {
  *n = g.n;
  *m = g.m;
  *R = (int32_t *)std::malloc(sizeof(int32_t) * (g.R.size() ? g.R.size() : 1));
  *C = (int32_t *)std::malloc(sizeof(int32_t) * (g.C.size() ? g.C.size() : 1));
  *F = (int32_t *)std::malloc(sizeof(int32_t) * (g.F.size() ? g.F.size() : 1));
  if (!g.R.empty()) std::memcpy(*R, g.R.data(), sizeof(int32_t) * g.R.size());
  if (!g.C.empty()) std::memcpy(*C, g.C.data(), sizeof(int32_t) * g.C.size());
  if (!g.F.empty()) std::memcpy(*F, g.F.data(), sizeof(int32_t) * g.F.size());
}

// The reference's parse_market / parse_snap (parse.cpp:210-434).  Arrays
// are returned with the lengths the reference's vectors have (C and F may
// be longer than m when the file holds more entries than its header).
int ref_parse_market(const char *path, int32_t *n, int64_t *m, int32_t **R, int32_t **C, int32_t **F,
                     int64_t *c_len) {
  std::vector<char> p(path, path + std::strlen(path) + 1);
  host_graph g = parse_market(p.data());
  copy_out(g, n, m, R, C, F);
  *c_len = (int64_t)g.C.size();
  return 0;
}

int ref_parse_snap(const char *path, int32_t header, int32_t *n, int64_t *m, int32_t **R, int32_t **C,
                   int32_t **F, int32_t *directed) {
  std::vector<char> p(path, path + std::strlen(path) + 1);
  host_graph g = parse_snap(p.data(), header != 0);
  copy_out(g, n, m, R, C, F);
  *directed = g.directed ? 1 : 0;
  return 0;
}

static host_graph make_graph(const int32_t *R, const int32_t *C, const int32_t *F, int32_t n, int64_t m,
                             int32_t directed) {
  host_graph g;
  g.n = n;
  g.m = (int)m;
  g.R.assign(R, R + n + 1);
  g.C.assign(C, C + m);
  g.F.assign(F, F + m);
  g.directed = directed != 0;
  return g;
}

// reverse (graph-utils/reverse.cpp:3-44).
int ref_reverse(const int32_t *R, const int32_t *C, const int32_t *F, int32_t n, int64_t m, int32_t **R2,
                int32_t **C2, int32_t **F2) {
  host_graph g = make_graph(R, C, F, n, m, 1);
  host_graph r = reverse(g);
  int32_t nn;
  int64_t mm;
  copy_out(r, &nn, &mm, R2, C2, F2);
  return 0;
}

// get_reachability_dag (graph-utils/dag_reachability.cpp:31-65) on a DAG;
// the sets flattened as offsets[n+1] / members.
int ref_reach_dag(const int32_t *R, const int32_t *C, const int32_t *F, int32_t n, int64_t m, int64_t **off,
                  int32_t **members) {
  host_graph g = make_graph(R, C, F, n, m, 1);
  std::vector<std::set<int>> reach = get_reachability_dag(g);
  *off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t total = 0;
  for (int i = 0; i < n; i++) total += (int64_t)reach[(size_t)i].size();
  *members = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
  int64_t k = 0;
  for (int i = 0; i < n; i++) {
    (*off)[i] = k;
    for (int v : reach[(size_t)i]) (*members)[k++] = v;
  }
  (*off)[n] = k;
  return 0;
}

void ref_free(void *p) { std::free(p); }
}
