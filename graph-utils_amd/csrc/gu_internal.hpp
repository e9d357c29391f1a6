// gu_internal.hpp -- library-internal entry points (not part of the C ABI).
//
// The row forms of gu_multi_search_rows / gu_betweenness_rows with one more
// host output, wrap[nsrc]: wrap[i] = 1 when a path count of source i wrapped
// modulo 2^64 (some reached vertex has at least 2^64 shortest paths).  The
// per-source engine reports it per source; the batched engine per launch
// (every source of a launch in which some source wrapped gets 1).  Used by
// the query server to flag each query's results.
#pragma once
#include "../../include/graph_utils_amd.h"

int gu_search_rows_wrap(gu_graph_t g, const int32_t *sources, int32_t nsrc, int32_t *const *dist_rows,
                        uint64_t *const *sigma_rows, int32_t *ecc, int32_t *wrap, const gu_config *cfg,
                        gu_stats *stats);
int gu_bc_rows_wrap(gu_graph_t g, const int32_t *sources, int32_t nsrc, float *const *delta_rows,
                    double *const *bc_targets, int32_t *wrap, const gu_config *cfg, gu_stats *stats);

// Allocates every workspace buffer that a launch of 1 to 256 sources of each
// job (search with or without path counts, Brandes, diameter) can need on
// handle g with configuration cfg, so that those launches allocate no device
// memory.  The query server calls it for each launch handle when a graph is
// registered: an allocation made while other streams run stalls their
// launches and copies, and one that has to wait for memory freed moments
// before took seconds.
int gu_reserve_launches(gu_graph_t g, const gu_config *cfg);

// A roctx range over a scope (gu_trace_push / gu_trace_pop): the engine's
// phases (level chunks, write-backs, Brandes backward, LBS) show up as
// ranges in `rocprofv3 --marker-trace`, with the kernels they launch.
struct GuTraceRange {
  explicit GuTraceRange(const char *m) : on(gu_trace_push(m) == GU_OK) {}
  ~GuTraceRange() {
    if (on) gu_trace_pop();
  }
  GuTraceRange(const GuTraceRange &) = delete;
  GuTraceRange &operator=(const GuTraceRange &) = delete;
  bool on;
};
