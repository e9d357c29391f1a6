// ref_parse_shim.cpp -- C entry points over the reference's own METIS reader.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's unmodified parse.cpp and util.cpp (read in place from the
// reference checkout) into oracle/_ref/libref_parse.so, so tests can compare
// the product's METIS reader with parse_metis (parse.cpp:115-208) itself.
// Nothing here restates the reader; it only copies host_graph's vectors out.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "parse.h"

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

void ref_free(void *p) { std::free(p); }
}
