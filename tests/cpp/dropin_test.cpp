// dropin_test.cpp -- a graph-utils example driver written against the
// reference's C++ signatures (graph_utils.hpp), checked with the CPU oracle.
// Mirrors examples/multi_search/multi_search_test.cu:129-277,
// examples/counting_shortest_paths, examples/betweenness_centrality and
// examples/diameter_sampling: parse -> device_graph -> *_setup -> verify the
// stored rows (row r holds the last source s = start + r mod grid).
//
// usage: dropin_test -i <graph.metis> -f dimacs [-k <sources>] [-n]
// Exit status 0 when every check passes.
#include <climits>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "../../graph-utils_amd/host/graph_utils.hpp"
#include "../../oracle/oracle.h"

static bool check_dist(const host_graph &g, const std::vector<std::vector<int>> &res, int start, int end,
                       const char *label) {
  dim3 grid, block;
  size_t rows = configure_grid(grid, block, start, end);
  std::vector<int32_t> src(rows);
  or_rows_for_range(start, end, (int)grid.x, src.data());
  if (res.size() != rows) {
    std::cout << label << ": wrong row count " << res.size() << " != " << rows << std::endl;
    return false;
  }
  std::vector<int32_t> d(g.n);
  for (size_t r = 0; r < rows; r++) {
    or_bfs(g.R.data(), g.C.data(), g.n, src[r], d.data());
    for (int v = 0; v < g.n; v++)
      if (d[v] != res[r][v]) {
        std::cout << "Mismatch for source " << src[r] << " and dest " << v << " (" << label << ")" << std::endl;
        return false;
      }
  }
  std::cout << label << ": Test passed." << std::endl;
  return true;
}

int main(int argc, char **argv) {
  program_options op;
  try {
    op = parse_arguments(argc, argv);
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 2;
  }
  host_graph g = parse(op);
  choose_device(op);
  device_graph g_d(g);
  int start = 0, end = op.approx ? std::min(op.approx, g.n) : g.n;
  std::cout << "graph n=" << g.n << " m=" << g.m << " sources=" << end - start << std::endl;
  bool ok = true;
  {
    // -n/--nvml: average power around the search (multi_search_test.cu:173-179)
    power_measurer pmes(10, op);
    if (op.nvml) pmes.start_power_sample();
    std::vector<std::vector<int>> r = multi_search_linear_atomics_setup(g_d, start, end);
    if (op.nvml)
      std::cout << "Average power: " << pmes.end_power_sample() << " W (energy counter: " << pmes.energy_average()
                << " W)" << std::endl;
    ok &= check_dist(g, r, start, end, "Linear with atomics");
  }
  ok &= check_dist(g, multi_search_edge_parallel_setup(g_d, start, end), start, end, "Edge parallel");
  ok &= check_dist(g, multi_search_warp_based_setup(g_d, start, end), start, end, "Warp based");
  ok &= check_dist(g, multi_search_scan_based_setup(g_d, start, end), start, end, "Scan based");
  ok &= check_dist(g, multi_search_CTA_warp_based_setup(g_d, start, end), start, end, "CTA+warp");
  ok &= check_dist(g, multi_search_shuffle_based_setup(g_d, op, start, end), start, end, "Shuffle based");
  {
    std::vector<std::vector<int>> tc;
    transitive_closure_setup(g_d, start, end, tc);  // examples/transitive_closure/reachability.cu
    ok &= check_dist(g, tc, start, end, "Transitive closure");
  }

  dim3 grid, block;
  size_t rows = configure_grid(grid, block, start, end);
  std::vector<int32_t> src(rows);
  or_rows_for_range(start, end, (int)grid.x, src.data());
  std::vector<int32_t> d(g.n);
  std::vector<uint64_t> s(g.n);
  std::vector<double> delta(g.n);

  std::vector<std::vector<unsigned long long>> sigma_h;
  count_shortest_paths_setup(g_d, start, end, sigma_h);
  bool sp = sigma_h.size() == rows;
  for (size_t r = 0; sp && r < rows; r++) {
    or_sigma(g.R.data(), g.C.data(), g.n, src[r], d.data(), s.data());
    for (int v = 0; v < g.n && sp; v++) sp = sigma_h[r][v] == s[v];
  }
  std::cout << "Shortest path counting: " << (sp ? "Test passed." : "Test failed.") << std::endl;
  ok &= sp;

  std::vector<std::vector<float>> delta_h;
  betweenness_centrality_setup(g_d, start, end, delta_h);
  double max_rel = 0, rms = 0;
  bool bc_ok = delta_h.size() == rows;
  for (size_t r = 0; bc_ok && r < rows; r++) {
    or_brandes_f64(g.R.data(), g.C.data(), g.n, src[r], nullptr, nullptr, delta.data());
    for (int v = 0; v < g.n; v++) {
      double e = std::fabs(delta[v] - (double)delta_h[r][v]);
      rms += e * e;
      double rel = delta[v] != 0 ? e / std::fabs(delta[v]) : (delta_h[r][v] == 0 ? 0.0 : 1.0);
      if (rel > max_rel) max_rel = rel;
    }
  }
  bc_ok = bc_ok && max_rel <= 1e-5;
  std::cout << "RMS Error: " << std::sqrt(rms / (double)g.n) << std::endl;
  std::cout << "Maximum relative error vs FP64: " << max_rel << std::endl;
  std::cout << "Betweenness centrality: " << (bc_ok ? "Test passed." : "Test failed.") << std::endl;
  ok &= bc_ok;

  int diam = diameter_sampling_setup(g_d, start, end);
  int expect = 0;
  for (int s0 = start; s0 < end; s0++) expect = std::max(expect, (int)or_eccentricity(g.R.data(), g.C.data(), g.n, s0));
  bool dok = diam == expect;
  std::cout << "Diameter sampling: " << diam << " (expected " << expect << ") "
            << (dok ? "Test passed." : "Test failed.") << std::endl;
  ok &= dok;
  return ok ? 0 : 1;
}
