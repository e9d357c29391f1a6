// reach_test.cpp -- the drop-in get_reachability_dag and reverse (reference
// graph-utils/dag_reachability.h, reverse.h) on random DAGs with hidden
// topological order, checked with the oracle's BFS (reachable = finite
// distance, not the source).  Exit status 0 when every check passes.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <random>
#include <vector>

#include "../../graph-utils_amd/host/graph_utils.hpp"
#include "../../oracle/oracle.h"

static host_graph random_dag(int n, int m, unsigned seed) {
  std::mt19937 rng(seed);
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::shuffle(perm.begin(), perm.end(), rng);
  std::vector<std::pair<int, int>> e;
  for (int k = 0; k < m; k++) {
    int u = (int)(rng() % n), v = (int)(rng() % n);
    if (u == v) continue;
    if (u > v) std::swap(u, v);
    e.emplace_back(perm[u], perm[v]);
  }
  std::sort(e.begin(), e.end());
  host_graph g;
  g.n = n;
  g.m = (int)e.size();
  g.R.assign(n + 1, 0);
  for (auto &p : e) {
    g.R[p.first + 1]++;
    g.C.push_back(p.second);
    g.F.push_back(p.first);
  }
  for (int v = 0; v < n; v++) g.R[v + 1] += g.R[v];
  g.directed = true;
  return g;
}

int main() {
  int fails = 0;
  for (unsigned seed = 1; seed <= 3; seed++) {
    const int n = seed == 3 ? 700 : 300;  // 700: three launches of 256 sources
    host_graph g = random_dag(n, 4 * n, seed);
    std::vector<std::set<int>> reach = get_reachability_dag(g);
    std::vector<int32_t> d(n);
    for (int s = 0; s < n; s++) {
      or_bfs(g.R.data(), g.C.data(), n, s, d.data());
      std::set<int> want;
      for (int v = 0; v < n; v++)
        if (d[v] != INT_MAX && d[v] > 0) want.insert(v);
      if (want != reach[(size_t)s]) {
        std::printf("reachability mismatch: seed %u source %d\n", seed, s);
        fails++;
        break;
      }
    }
    host_graph r = reverse(g);
    host_graph rr = reverse(r);
    if (rr.R != g.R || rr.C != g.C) {
      std::printf("reverse twice differs (seed %u)\n", seed);
      fails++;
    }
  }
  host_graph u = random_dag(50, 100, 9);
  u.directed = false;
  bool threw = false;
  try {
    get_reachability_dag(u);
  } catch (const gu_error &e) {
    threw = e.status() == GU_ERR_INVALID;
  }
  if (!threw) {
    std::printf("undirected input not refused\n");
    fails++;
  }
  std::printf("reach_test: %s\n", fails ? "Test failed." : "Test passed.");
  return fails ? 1 : 0;
}
