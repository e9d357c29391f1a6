// graph_utils.hpp -- the reference's C++ host interface for the multi-search
// path, kept signature-for-signature so drivers written against graph-utils
// compile unchanged, implemented over the C ABI (include/graph_utils_amd.h).
//
//   program_options, parse_arguments      util.h:10-25, util.cpp:3-77
//   host_graph, parse, parse_metis,       parse.h:15-40, parse.cpp:88-456
//   parse_snap, parse_market,
//   write_edgelist_to_file
//   device_graph                          device_graph.h:7-22
//   start_clock, end_clock, choose_device util_device.cuh:38-41
//   power_measurer                        util_device.cuh:47-62, util_device.cu:79-130
//   configure_grid, transfer_result       common.cuh:9-14
//   multi_search_*_setup                  linear_with_atomics.cuh:14,
//                                         edge_parallel.cuh:13, warp_based.cuh:15,
//                                         scan_based.cuh:15, CTA_warp.cuh:15,
//                                         shuffle_based.cuh:16
//   count_shortest_paths_setup            count_shortest_paths.cuh:13
//   transitive_closure_setup              transitive_closure.cuh:13
//   betweenness_centrality_setup          betweenness_centrality.cuh:14
//   diameter_sampling_setup               diameter_sampling.cuh:13
//   reverse, get_reachability_dag         graph-utils/reverse.h, dag_reachability.h
//
// Differences in error behaviour: where the reference prints and calls exit()
// (checkCudaErrors, parse errors) these functions throw gu_error carrying the
// C ABI status; nothing here terminates the process.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/graph_utils_amd.h"

class gu_error : public std::runtime_error {
 public:
  gu_error(int status, const std::string &what)
      : std::runtime_error(what + ": " + gu_strerror(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

class program_options {
 public:
  program_options()
      : device(-1), infile(nullptr), outfile(nullptr), verify(false), approx(0), format(nullptr),
        isTesla(false), nvml(false) {}
  int device;
  char *infile;
  char *outfile;
  bool verify;
  int approx;
  char *format;
  bool isTesla;  // kept for layout compatibility; power_measurer asks AMD SMI instead
  bool nvml;
};

// Flags -d/--device -f/--format -h/--help -k/--approx -i/--infile -n/--nvml
// -o/--outfile -v/--verify.  Throws std::invalid_argument on bad input.
program_options parse_arguments(int argc, char *argv[]);

class host_graph {
 public:
  void print_offset_array() const;
  void print_edge_array() const;
  void print_from_array() const;
  void print_adjacency_list() const;
  unsigned long long count_degree_zero_vertices() const;
  bool write_edgelist_to_file(const std::string &file, bool header) const;  // true on success

  std::vector<int> C;  // column ids
  std::vector<int> R;  // CSR offsets
  std::vector<int> F;  // row of each entry
  int n = 0;
  int m = 0;  // directed entries (an undirected edge counts twice)
  bool directed = false;
};

host_graph parse(const program_options &op);  // dimacs, edgelist, edgelist_h, market
host_graph parse_metis(char *file);
host_graph parse_snap(char *file, bool header);
host_graph parse_market(char *file);

class device_graph {
 public:
  explicit device_graph(const host_graph &g);
  ~device_graph();
  device_graph(const device_graph &) = delete;
  device_graph &operator=(const device_graph &) = delete;

  const int *R = nullptr;  // device pointers
  const int *C = nullptr;
  const int *F = nullptr;
  int n = 0;
  int m = 0;
  bool directed = false;
  gu_graph_t handle = nullptr;
};

void start_clock(hipEvent_t &start, hipEvent_t &end);
float end_clock(hipEvent_t &start, hipEvent_t &end);  // seconds
void choose_device(program_options &op);

// Samples the chosen device's socket power through AMD SMI every `p` ms
// between start_power_sample() and end_power_sample(), on a host thread.
// end_power_sample() synchronizes the device and returns the mean of the
// samples in W, or -1 (after a warning at start) where the board's power
// cannot be read, as the reference does on boards other than Tesla.
class power_measurer {
 public:
  power_measurer(long p, const program_options &op) : period(p), device(op.device < 0 ? 0 : op.device) {}
  ~power_measurer();
  power_measurer(const power_measurer &) = delete;
  power_measurer &operator=(const power_measurer &) = delete;
  void start_power_sample();
  double end_power_sample();
  // energy used between start and end divided by the interval, W (-1 when
  // the board has no energy counter); valid after end_power_sample()
  double energy_average() const { return energy_avg; }

 private:
  power_measurer();
  long period;
  int device;
  gu_power_t handle = nullptr;
  double energy_avg = -1.0;
};

// Grid = compute units, block = max threads per block; returns the number of
// result rows min(end - start, compute units).
size_t configure_grid(dim3 &dimGrid, dim3 &dimBlock, int start, int end);

void transfer_result(const device_graph &g, int *d_d, size_t pitch_d, size_t sources_to_store,
                     std::vector<std::vector<int>> &d_host_vector);
void transfer_result(const device_graph &g, unsigned long long *d_d, size_t pitch_d,
                     size_t sources_to_store, std::vector<std::vector<unsigned long long>> &d_host_vector);
void transfer_result(const device_graph &g, float *d_d, size_t pitch_d, size_t sources_to_store,
                     std::vector<std::vector<float>> &d_host_vector);

// Every source in [start,end) is searched; row r of the result holds the
// last source s with s = start + r (mod compute units), as the reference's
// one-block-per-source kernels leave it.
std::vector<std::vector<int>> multi_search_linear_atomics_setup(const device_graph &g, int start, int end);
std::vector<std::vector<int>> multi_search_edge_parallel_setup(const device_graph &g, int start, int end);
std::vector<std::vector<int>> multi_search_warp_based_setup(const device_graph &g, int start, int end);
std::vector<std::vector<int>> multi_search_scan_based_setup(const device_graph &g, int start, int end);
std::vector<std::vector<int>> multi_search_CTA_warp_based_setup(const device_graph &g, int start, int end);
std::vector<std::vector<int>> multi_search_shuffle_based_setup(const device_graph &g, const program_options &op,
                                                               int start, int end);
void count_shortest_paths_setup(const device_graph &g, int start, int end,
                                std::vector<std::vector<unsigned long long>> &sigma_h);
// transitive_closure.cuh:13 -- distance rows; INT_MAX marks unreachable.
void transitive_closure_setup(const device_graph &g, int start, int end, std::vector<std::vector<int>> &d);
void betweenness_centrality_setup(const device_graph &g, int start, int end,
                                  std::vector<std::vector<float>> &delta_h);
int diameter_sampling_setup(const device_graph &g, int start, int end);

// reverse (graph-utils/reverse.h): the transposed graph, entries ordered by
// (target, source).
host_graph reverse(host_graph &g);
// get_reachability_dag (graph-utils/dag_reachability.h): for every vertex of
// a directed acyclic graph, the set of vertices reachable from it by at
// least one edge.  Computed with multi-source searches on the GPU (256
// sources per launch); throws gu_error for an undirected graph, as the
// reference exits.
std::vector<std::set<int>> get_reachability_dag(host_graph &g);

// Rows of a [start,end) result and the source each row holds.
std::vector<int> result_row_sources(int start, int end, int grid);
