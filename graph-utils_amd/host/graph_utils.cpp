// graph_utils.cpp -- the reference's host interface over the C ABI.
#include "graph_utils.hpp"

#include <getopt.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>

namespace {

void check(int rc, const char *what) {
  if (rc != GU_OK) throw gu_error(rc, what);
}
void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) throw gu_error(GU_ERR_HIP, std::string(what) + " (" + hipGetErrorString(e) + ")");
}

bool verbose() {
  const char *v = std::getenv("GU_VERBOSE");
  return v && *v && *v != '0';
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  explicit DevBuf(size_t count) { check(gu_malloc((void **)&p, sizeof(T) * (count ? count : 1)), "gu_malloc"); }
  ~DevBuf() {
    if (p) gu_free(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

template <class T>
void transfer_rows(const device_graph &g, T *d, size_t pitch, size_t rows, std::vector<std::vector<T>> &out) {
  std::vector<T> flat((size_t)g.n * rows);
  if (rows && g.n)
    check_hip(hipMemcpy2D(flat.data(), sizeof(T) * g.n, d, pitch, sizeof(T) * g.n, rows, hipMemcpyDeviceToHost),
              "transfer_result");
  out.assign(rows, std::vector<T>());
  for (size_t i = 0; i < rows; i++) out[i].assign(flat.begin() + i * g.n, flat.begin() + (i + 1) * g.n);
}

// Stored-row sources first (their rows are kept), then every other source of
// [start,end) (searched, nothing kept), as the reference's kernels do.
void split_sources(int start, int end, int grid, std::vector<int> &stored, std::vector<int> &rest) {
  stored = result_row_sources(start, end, grid);
  std::vector<char> keep((size_t)(end - start), 0);
  for (int s : stored) keep[(size_t)(s - start)] = 1;
  rest.clear();
  for (int s = start; s < end; s++)
    if (!keep[(size_t)(s - start)]) rest.push_back(s);
}

std::vector<std::vector<int>> run_distances(const device_graph &g, int start, int end, int variant,
                                            const char *label) {
  std::vector<std::vector<int>> out;
  if (start < 0 || end > g.n || end < start) throw gu_error(GU_ERR_RANGE, "source range");
  dim3 grid, block;
  size_t rows = configure_grid(grid, block, start, end);
  if (rows == 0) return out;
  std::vector<int> stored, rest;
  split_sources(start, end, (int)grid.x, stored, rest);
  hipEvent_t a, b;
  start_clock(a, b);
  DevBuf<int> d(rows * (size_t)g.n);
  gu_config cfg{};
  cfg.variant = variant;
  check(gu_multi_search(g.handle, stored.data(), (int)stored.size(), d.p, nullptr, &cfg, nullptr), label);
  if (!rest.empty())
    check(gu_multi_search(g.handle, rest.data(), (int)rest.size(), nullptr, nullptr, &cfg, nullptr), label);
  transfer_result(g, d.p, sizeof(int) * g.n, rows, out);
  float t = end_clock(a, b);
  if (verbose()) std::cout << "Time for " << label << ": " << std::setprecision(9) << t << " s" << std::endl;
  return out;
}

}  // namespace

program_options parse_arguments(int argc, char *argv[]) {
  program_options op;
  static struct option long_options[] = {{"device", required_argument, 0, 'd'},
                                         {"format", required_argument, 0, 'f'},
                                         {"help", no_argument, 0, 'h'},
                                         {"approx", required_argument, 0, 'k'},
                                         {"infile", required_argument, 0, 'i'},
                                         {"nvml", no_argument, 0, 'n'},
                                         {"outfile", required_argument, 0, 'o'},
                                         {"verify", no_argument, 0, 'v'},
                                         {0, 0, 0, 0}};
  int c, idx = 0;
  optind = 1;
  while ((c = getopt_long(argc, argv, "d:f:hk:i:no:v", long_options, &idx)) != -1) {
    switch (c) {
      case 'd': op.device = std::atoi(optarg); break;
      case 'f': op.format = optarg; break;
      case 'h':
        throw std::invalid_argument(std::string("Usage: ") + argv[0] +
                                    " -i <input graph file> -f <graph format: dimacs>");
      case 'k': op.approx = std::atoi(optarg); break;
      case 'i': op.infile = optarg; break;
      case 'n': op.nvml = true; break;
      case 'o': op.outfile = optarg; break;
      case 'v': op.verify = true; break;
      default: throw std::invalid_argument("bad command line argument");
    }
  }
  if (!op.infile || !op.format)
    throw std::invalid_argument("Input graph file and format are required. Use the -i and -f switches.");
  return op;
}

void host_graph::print_offset_array() const {
  std::cout << "R = [";
  for (size_t i = 0; i < R.size(); i++) std::cout << (i ? "," : "") << R[i];
  std::cout << "]" << std::endl;
}
void host_graph::print_edge_array() const {
  std::cout << "C = [";
  for (size_t i = 0; i < C.size(); i++) std::cout << (i ? "," : "") << C[i];
  std::cout << "]" << std::endl;
}
void host_graph::print_from_array() const {
  std::cout << "F = [";
  for (size_t i = 0; i < F.size(); i++) std::cout << (i ? "," : "") << F[i];
  std::cout << "]" << std::endl;
}
void host_graph::print_adjacency_list() const {
  std::cout << "Edge lists for each vertex: " << std::endl;
  for (int i = 0; i < n; i++) {
    std::cout << i << " | ";
    for (int j = R[i]; j < R[i + 1]; j++) std::cout << (j > R[i] ? ", " : "") << C[j];
    std::cout << std::endl;
  }
}
unsigned long long host_graph::count_degree_zero_vertices() const {
  unsigned long long z = 0;
  for (int i = 0; i < n; i++) z += R[i + 1] == R[i];
  return z;
}

namespace {
host_graph take_graph(int32_t *R, int32_t *C, int32_t *F, int32_t n, int64_t m, bool directed) {
  host_graph g;
  g.n = n;
  g.m = (int)m;
  g.R.assign(R, R + n + 1);
  g.C.assign(C, C + m);
  g.F.assign(F, F + m);
  g.directed = directed;
  gu_host_free(R);
  gu_host_free(C);
  gu_host_free(F);
  return g;
}
}  // namespace

host_graph parse_metis(char *file) {
  int32_t *R = nullptr, *C = nullptr, *F = nullptr, n = 0;
  int64_t m = 0;
  check(gu_parse_metis(file, &R, &C, &F, &n, &m), "parse_metis");
  return take_graph(R, C, F, n, m, false);
}

host_graph parse_market(char *file) {
  int32_t *R = nullptr, *C = nullptr, *F = nullptr, n = 0;
  int64_t m = 0;
  check(gu_parse_market(file, &R, &C, &F, &n, &m), "parse_market");
  return take_graph(R, C, F, n, m, false);
}

host_graph parse_snap(char *file, bool header) {
  int32_t *R = nullptr, *C = nullptr, *F = nullptr, n = 0;
  int64_t m = 0;
  check(gu_parse_snap(file, header ? 1 : 0, &R, &C, &F, &n, &m), "parse_snap");
  return take_graph(R, C, F, n, m, true);
}

bool host_graph::write_edgelist_to_file(const std::string &file, bool header) const {
  return gu_write_edgelist(file.c_str(), header ? 1 : 0, F.data(), C.data(), n, m) == GU_OK;
}

// parse (parse.cpp:88-113): dimacs (METIS), edgelist, edgelist_h, market.
host_graph parse(const program_options &op) {
  const std::string f = op.format ? op.format : "";
  if (f == "dimacs") return parse_metis(op.infile);
  if (f == "edgelist") return parse_snap(op.infile, false);
  if (f == "edgelist_h") return parse_snap(op.infile, true);
  if (f == "market") return parse_market(op.infile);
  throw gu_error(GU_ERR_INVALID, "unsupported graph file type");
}

device_graph::device_graph(const host_graph &g) {
  check(gu_graph_create(g.R.data(), g.C.data(), g.F.empty() ? nullptr : g.F.data(), g.n, g.m,
                        g.directed ? 1 : 0, &handle),
        "device_graph");
  const int32_t *r, *c, *f;
  int32_t nn;
  int64_t mm;
  check(gu_graph_arrays(handle, &r, &c, &f, &nn, &mm), "device_graph");
  R = r;
  C = c;
  F = f;
  n = nn;
  m = (int)mm;
  directed = g.directed;
}

device_graph::~device_graph() {
  if (handle) gu_graph_destroy(handle);
}

void start_clock(hipEvent_t &start, hipEvent_t &end) {
  check_hip(hipEventCreate(&start), "start_clock");
  check_hip(hipEventCreate(&end), "start_clock");
  check_hip(hipEventRecord(start, 0), "start_clock");
}

float end_clock(hipEvent_t &start, hipEvent_t &end) {
  float ms = 0;
  check_hip(hipEventRecord(end, 0), "end_clock");
  check_hip(hipEventSynchronize(end), "end_clock");
  check_hip(hipEventElapsedTime(&ms, start, end), "end_clock");
  check_hip(hipEventDestroy(start), "end_clock");
  check_hip(hipEventDestroy(end), "end_clock");
  return ms / 1000.0f;
}

void choose_device(program_options &op) {
  gu_device_props p;
  check(gu_choose_device(op.device, &p), "choose_device");
  op.device = p.device;
  op.isTesla = false;
  if (verbose()) {
    std::cout << "Chosen Device: " << p.name << " (" << p.arch << ")" << std::endl;
    std::cout << "Number of Compute Units: " << p.compute_units << std::endl;
    std::cout << "Size of Global Memory: " << p.total_mem / (double)(1 << 30) << " GB" << std::endl;
    std::cout << "Memory Bandwidth: " << p.mem_bandwidth_gbs << " GB/s " << std::endl;
  }
}

power_measurer::~power_measurer() {
  if (handle) gu_power_stop(handle, nullptr);
}

void power_measurer::start_power_sample() {
  if (handle) return;
  const int32_t ms = period < 1 ? 1 : (period > 60000 ? 60000 : (int32_t)period);
  const int rc = gu_power_start(device, ms, &handle);
  if (rc == GU_ERR_UNSUPPORTED) {
    handle = nullptr;
    std::cerr << "Warning: the power of this device cannot be read (AMD SMI)." << std::endl;
    return;
  }
  check(rc, "start_power_sample");
}

double power_measurer::end_power_sample() {
  if (!handle) return -1;
  gu_power_result r;
  gu_power_t h = handle;
  handle = nullptr;
  check(gu_power_stop(h, &r), "end_power_sample");
  energy_avg = r.energy_avg_w;
  return r.avg_w;
}

size_t configure_grid(dim3 &dimGrid, dim3 &dimBlock, int start, int end) {
  int dev = 0;
  check_hip(hipGetDevice(&dev), "configure_grid");
  hipDeviceProp_t p;
  check_hip(hipGetDeviceProperties(&p, dev), "configure_grid");
  dimGrid = dim3(p.multiProcessorCount, 1, 1);
  dimBlock = dim3(p.maxThreadsPerBlock, 1, 1);
  if (end <= start) return 0;
  return (size_t)(end - start) < dimGrid.x ? (size_t)(end - start) : dimGrid.x;
}

std::vector<int> result_row_sources(int start, int end, int grid) {
  std::vector<int> rows;
  const int total = end - start;
  if (total <= 0 || grid <= 0) return rows;
  const int r = total < grid ? total : grid;
  rows.resize(r);
  if (total <= grid) {
    for (int i = 0; i < r; i++) rows[i] = start + i;
  } else {
    // The last source of block b is the largest s < end with s = start + b
    // (mod grid): the sources end-grid .. end-1 land on rows (total+i) % grid.
    for (int i = 0; i < r; i++) rows[(total + i) % r] = end - r + i;
  }
  return rows;
}

void transfer_result(const device_graph &g, int *d_d, size_t pitch_d, size_t rows,
                     std::vector<std::vector<int>> &out) {
  transfer_rows(g, d_d, pitch_d, rows, out);
}
void transfer_result(const device_graph &g, unsigned long long *d_d, size_t pitch_d, size_t rows,
                     std::vector<std::vector<unsigned long long>> &out) {
  transfer_rows(g, d_d, pitch_d, rows, out);
}
void transfer_result(const device_graph &g, float *d_d, size_t pitch_d, size_t rows,
                     std::vector<std::vector<float>> &out) {
  transfer_rows(g, d_d, pitch_d, rows, out);
}

std::vector<std::vector<int>> multi_search_linear_atomics_setup(const device_graph &g, int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_LINEAR, "baseline linear algorithm with atomics");
}
std::vector<std::vector<int>> multi_search_edge_parallel_setup(const device_graph &g, int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_EDGE_PARALLEL, "edge-parallel algorithm");
}
std::vector<std::vector<int>> multi_search_warp_based_setup(const device_graph &g, int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_WARP, "warp-based neighbor gathering");
}
std::vector<std::vector<int>> multi_search_scan_based_setup(const device_graph &g, int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_SCAN, "scan-based neighbor gathering");
}
std::vector<std::vector<int>> multi_search_CTA_warp_based_setup(const device_graph &g, int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_CTA_WARP, "CTA+warp neighbor gathering");
}
std::vector<std::vector<int>> multi_search_shuffle_based_setup(const device_graph &g, const program_options &,
                                                               int start, int end) {
  return run_distances(g, start, end, GU_VARIANT_SHUFFLE, "shuffle-based neighbor gathering");
}

host_graph reverse(host_graph &g) {
  int32_t *R = nullptr, *C = nullptr, *F = nullptr;
  check(gu_reverse(g.R.data(), g.C.data(), g.n, g.m, &R, &C, &F), "reverse");
  return take_graph(R, C, F, g.n, g.m, g.directed);
}

std::vector<std::set<int>> get_reachability_dag(host_graph &g) {
  if (!g.directed) throw gu_error(GU_ERR_INVALID, "reachability needs a directed graph");
  std::vector<std::set<int>> reach((size_t)g.n);
  if (g.n == 0) return reach;
  gu_graph_t h = nullptr;
  check(gu_graph_create(g.R.data(), g.C.data(), g.F.data(), g.n, g.m, 1, &h), "gu_graph_create");
  const int chunk = std::min(g.n, 256);
  int32_t *dist = nullptr;
  int rc = gu_malloc((void **)&dist, sizeof(int32_t) * (size_t)chunk * g.n);
  std::vector<int32_t> rows((size_t)chunk * g.n), src((size_t)chunk);
  for (int b0 = 0; rc == GU_OK && b0 < g.n; b0 += chunk) {
    const int nb = std::min(chunk, g.n - b0);
    for (int i = 0; i < nb; i++) src[(size_t)i] = b0 + i;
    rc = gu_multi_search(h, src.data(), nb, dist, nullptr, nullptr, nullptr);
    if (rc == GU_OK) rc = gu_memcpy_d2h(rows.data(), dist, sizeof(int32_t) * (size_t)nb * g.n);
    for (int i = 0; rc == GU_OK && i < nb; i++) {
      std::set<int> &s = reach[(size_t)(b0 + i)];
      const int32_t *d = rows.data() + (size_t)i * g.n;
      for (int v = 0; v < g.n; v++)
        if (d[v] != INT_MAX && d[v] > 0) s.insert(s.end(), v);
    }
  }
  if (dist) gu_free(dist);
  gu_graph_destroy(h);
  check(rc, "get_reachability_dag");
  return reach;
}

// transitive_closure_setup (transitive_closure.cu:4-37): the reachability
// rows are the BFS distance rows (INT_MAX = not reachable).
void transitive_closure_setup(const device_graph &g, int start, int end, std::vector<std::vector<int>> &d) {
  d = run_distances(g, start, end, GU_VARIANT_AUTO, "shuffle-based transitive closure");
}

void count_shortest_paths_setup(const device_graph &g, int start, int end,
                                std::vector<std::vector<unsigned long long>> &sigma_h) {
  sigma_h.clear();
  if (start < 0 || end > g.n || end < start) throw gu_error(GU_ERR_RANGE, "source range");
  dim3 grid, block;
  size_t rows = configure_grid(grid, block, start, end);
  if (rows == 0) return;
  std::vector<int> stored, rest;
  split_sources(start, end, (int)grid.x, stored, rest);
  hipEvent_t a, b;
  start_clock(a, b);
  DevBuf<uint64_t> s(rows * (size_t)g.n);
  gu_config cfg{};
  check(gu_multi_search(g.handle, stored.data(), (int)stored.size(), nullptr, s.p, &cfg, nullptr),
        "count_shortest_paths");
  if (!rest.empty()) {
    DevBuf<uint64_t> scratch((size_t)std::min<size_t>(rest.size(), 256) * g.n);
    for (size_t i = 0; i < rest.size(); i += 256) {
      int k = (int)std::min<size_t>(256, rest.size() - i);
      check(gu_multi_search(g.handle, rest.data() + i, k, nullptr, scratch.p, &cfg, nullptr),
            "count_shortest_paths");
    }
  }
  transfer_result(g, (unsigned long long *)s.p, sizeof(uint64_t) * g.n, rows, sigma_h);
  float t = end_clock(a, b);
  if (verbose())
    std::cout << "Time for shuffle-based shortest-path counting: " << std::setprecision(9) << t << " s" << std::endl;
}

void betweenness_centrality_setup(const device_graph &g, int start, int end,
                                  std::vector<std::vector<float>> &delta_h) {
  delta_h.clear();
  if (start < 0 || end > g.n || end < start) throw gu_error(GU_ERR_RANGE, "source range");
  dim3 grid, block;
  size_t rows = configure_grid(grid, block, start, end);
  if (rows == 0) return;
  std::vector<int> stored, rest;
  split_sources(start, end, (int)grid.x, stored, rest);
  hipEvent_t a, b;
  start_clock(a, b);
  DevBuf<float> d(rows * (size_t)g.n);
  DevBuf<double> bc((size_t)g.n);
  check(gu_memset_d(bc.p, 0, sizeof(double) * g.n), "betweenness_centrality");
  gu_config cfg{};
  gu_stats st{};
  check(gu_betweenness(g.handle, stored.data(), (int)stored.size(), d.p, bc.p, &cfg, &st),
        "betweenness_centrality");
  int depth = st.max_level;
  if (!rest.empty()) {
    check(gu_betweenness(g.handle, rest.data(), (int)rest.size(), nullptr, bc.p, &cfg, &st),
          "betweenness_centrality");
    depth = std::max(depth, st.max_level);
  }
  transfer_result(g, d.p, sizeof(float) * g.n, rows, delta_h);
  float t = end_clock(a, b);
  if (verbose()) {
    std::cout << "Estimated diameter: " << depth << std::endl;
    std::cout << "Time for shuffle-based BC: " << std::setprecision(9) << t << " s" << std::endl;
  }
}

int diameter_sampling_setup(const device_graph &g, int start, int end) {
  if (start < 0 || end > g.n || end < start) throw gu_error(GU_ERR_RANGE, "source range");
  if (end == start) return 0;
  std::vector<int> src;
  for (int s = start; s < end; s++) src.push_back(s);
  hipEvent_t a, b;
  start_clock(a, b);
  int32_t diam = 0;
  gu_config cfg{};
  check(gu_diameter(g.handle, src.data(), (int)src.size(), &diam, &cfg, nullptr), "diameter_sampling");
  float t = end_clock(a, b);
  if (verbose())
    std::cout << "Time for shuffle-based diameter sampling: " << std::setprecision(9) << t << " s" << std::endl;
  return diam;
}
