// gu_host.cpp -- host-side graph I/O and seeded generators (plain C++17).
//
// gu_parse_metis restates parse_metis (parse.cpp:115-208) token for token,
// including how it splits lines: a blank arriving while the token buffer is
// empty is kept in the token (parse.cpp:143-154), so a leading or doubled
// blank makes the next token non-numeric and it is dropped with R adjusted
// (parse.cpp:184-193).  Where the reference would write out of bounds
// (more adjacency lines than n, more entries than the header's m) or store
// an id outside [0,n), this reader returns an error instead.
//
// Generators use a project RNG (splitmix64 seeding, xoshiro256** streams),
// so every graph is a pure function of its parameters and seed and does not
// depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/graph_utils_amd.h"

namespace {

inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t mix64(uint64_t x) {
  uint64_t s = x;
  return splitmix64(s);
}

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // Uniform integer in [0, bound) by rejection (no modulo bias).
  inline uint64_t below(uint64_t bound) {
    if (bound <= 1) return 0;
    uint64_t lim = (~0ull) - ((~0ull) % bound);
    uint64_t x;
    do x = next(); while (x >= lim);
    return x % bound;
  }
};

int n_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  if (h > 16) h = 16;
  return (int)h;
}

template <class F>
void parallel_for(int64_t n, F f) {
  int t = n_threads();
  if (n < 65536 || t == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int i = 0; i < t; i++) {
    int64_t lo = n * i / t, hi = n * (i + 1) / t;
    th.emplace_back([=]() { f(lo, hi); });
  }
  for (auto &x : th) x.join();
}

template <class T>
T *dup_out(const std::vector<T> &v) {
  T *p = (T *)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

// Undirected edge list -> symmetric CSR with sorted, duplicate-free rows and
// no self loops.  F[k] is the row of entry k.
int build_csr(int32_t n, const std::vector<int32_t> &eu, const std::vector<int32_t> &ev,
              int32_t **R, int32_t **C, int32_t **F, int32_t *n_out, int64_t *m_out) {
  const int64_t E = (int64_t)eu.size();
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t i = 0; i < E; i++) {
    if (eu[i] == ev[i]) continue;
    deg[eu[i]]++;
    deg[ev[i]]++;
  }
  std::vector<int64_t> off(n + 1, 0);
  for (int32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
  std::vector<int32_t> adj(off[n]);
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (int64_t i = 0; i < E; i++) {
    int32_t u = eu[i], v = ev[i];
    if (u == v) continue;
    adj[pos[u]++] = v;
    adj[pos[v]++] = u;
  }
  std::vector<int64_t> nd(n, 0);
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t v = lo; v < hi; v++) {
      auto b = adj.begin() + off[v], e = adj.begin() + off[v + 1];
      std::sort(b, e);
      nd[v] = std::unique(b, e) - b;
    }
  });
  std::vector<int32_t> Rv(n + 1, 0);
  int64_t tot = 0;
  for (int32_t v = 0; v < n; v++) {
    Rv[v] = (int32_t)tot;
    tot += nd[v];
  }
  if (tot >= (int64_t)INT32_MAX) return GU_ERR_CAPACITY;
  Rv[n] = (int32_t)tot;
  std::vector<int32_t> Cv(tot), Fv(tot);
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t v = lo; v < hi; v++) {
      std::memcpy(Cv.data() + Rv[v], adj.data() + off[v], sizeof(int32_t) * nd[v]);
      for (int64_t k = Rv[v]; k < Rv[v + 1]; k++) Fv[k] = (int32_t)v;
    }
  });
  *R = dup_out(Rv);
  *C = dup_out(Cv);
  if (F) *F = dup_out(Fv);
  *n_out = n;
  *m_out = tot;
  return (*R && *C) ? GU_OK : GU_ERR_NOMEM;
}

bool is_number(const std::string &s) {
  if (s.empty()) return false;
  for (char ch : s)
    if (!(ch >= '0' && ch <= '9')) return false;
  return true;
}

bool to_int(const std::string &s, long long *out) {
  // std::stoi semantics: leading blanks skipped, optional sign, digits.
  const char *p = s.c_str();
  char *end = nullptr;
  errno = 0;
  long long v = std::strtoll(p, &end, 10);
  if (end == p || errno != 0 || v > INT32_MAX || v < INT32_MIN) return false;
  *out = v;
  return true;
}

// The reference readers' line splitter (parse.cpp:138-157 and its copies in
// parse_market / parse_snap): a blank or tab ends a non-empty token; one
// arriving while the token is empty is kept in it.
void split_line(const std::string &line, std::vector<std::string> &tok) {
  tok.clear();
  std::string t;
  for (char ch : line) {
    if ((ch == '\t' || ch == ' ') && !t.empty()) {
      tok.push_back(t);
      t.clear();
    } else {
      t.push_back(ch);
    }
  }
  if (!t.empty()) tok.push_back(t);
}

// R from the row of every entry, as the reference readers induce it
// (parse.cpp:319-331, 416-429): R[v] is the index of the first entry whose
// row exceeds v-1 while rows ascend; rows that never appear take the
// position reached.  Rows must lie in [0, n).
void induce_R(int64_t n, const std::vector<int32_t> &Fv, int64_t m, std::vector<int32_t> &Rv) {
  Rv.assign((size_t)n + 1, 0);
  int64_t last = 0;
  for (int64_t i = 0; i < m; i++)
    while (Fv[(size_t)i] > last && last < n + 1) Rv[(size_t)++last] = (int32_t)i;
  while (last < n) Rv[(size_t)++last] = (int32_t)m;
}

}  // namespace

extern "C" {

void gu_host_free(void *p) { std::free(p); }

// parse_market (parse.cpp:210-332): pattern Matrix Market; '%' lines
// skipped; the first other line holds rows, columns and nonzeros (m = 2 x
// nonzeros); every further line "u v" (1-based) adds (u,v) and (v,u); the
// entries are sorted by (row, column) and R induced from the rows.  Where
// the reference prints and exits (wrong field counts) or reads past its
// arrays (fewer entries than the header, ids outside [1,n]) this reader
// returns an error; more entries than the header keep the first m after
// sorting, as the reference's R does.
int gu_parse_market(const char *path, int32_t **R, int32_t **C, int32_t **F, int32_t *n_out,
                    int64_t *m_out) {
  if (!path || !R || !C || !n_out || !m_out) return GU_ERR_INVALID;
  std::ifstream in(path, std::ifstream::in);
  if (!in.good()) return GU_ERR_IO;
  std::string line;
  std::vector<std::string> tok;
  bool first = true;
  long long n = 0, m = 0;
  std::vector<std::pair<int32_t, int32_t>> e;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '%') continue;
    split_line(line, tok);
    if (first) {
      long long a, b, c;
      if (tok.size() != 3) return GU_ERR_PARSE;
      if (!to_int(tok[0], &a) || !to_int(tok[1], &b) || !to_int(tok[2], &c)) return GU_ERR_PARSE;
      if (a < 0 || c < 0 || 2 * c >= INT32_MAX) return GU_ERR_PARSE;
      n = a;
      m = 2 * c;  // a non-square header is only warned about by the reference
      e.reserve((size_t)m);
      first = false;
      continue;
    }
    if (tok.size() != 2) return GU_ERR_PARSE;  // pattern matrices only
    long long u, v;
    if (!to_int(tok[0], &u) || !to_int(tok[1], &v)) return GU_ERR_PARSE;
    if (u < 1 || u > n || v < 1 || v > n) return GU_ERR_RANGE;
    e.emplace_back((int32_t)(u - 1), (int32_t)(v - 1));
    e.emplace_back((int32_t)(v - 1), (int32_t)(u - 1));
  }
  if (first) return GU_ERR_PARSE;
  if ((long long)e.size() < m) return GU_ERR_PARSE;
  std::sort(e.begin(), e.end());
  std::vector<int32_t> Fv((size_t)m), Cv((size_t)m), Rv;
  for (long long i = 0; i < m; i++) {
    Fv[(size_t)i] = e[(size_t)i].first;
    Cv[(size_t)i] = e[(size_t)i].second;
  }
  induce_R(n, Fv, m, Rv);
  *R = dup_out(Rv);
  *C = dup_out(Cv);
  if (F) *F = dup_out(Fv);
  *n_out = (int32_t)n;
  *m_out = m;
  return (*R && *C) ? GU_OK : GU_ERR_NOMEM;
}

// parse_snap (parse.cpp:335-434): '#' lines skipped; each other line "u v"
// (0-based, further fields ignored) is one directed entry in file order;
// n = number of distinct ids, m = number of lines, R induced from the rows
// in file order (a file not sorted by source yields the reference's R for
// it).  The reference ignores its header flag, so a header line of an
// "edgelist_h" file is read as an entry; `header` is accepted and ignored
// the same way.  Where the reference reads out of bounds (fewer than two
// fields, ids outside [0,n)) this reader returns an error.  The graph is
// directed.
int gu_parse_snap(const char *path, int32_t header, int32_t **R, int32_t **C, int32_t **F,
                  int32_t *n_out, int64_t *m_out) {
  (void)header;
  if (!path || !R || !C || !n_out || !m_out) return GU_ERR_INVALID;
  std::ifstream in(path, std::ifstream::in);
  if (!in.good()) return GU_ERR_IO;
  std::string line;
  std::vector<std::string> tok;
  std::vector<int32_t> Fv, Cv, Rv;
  std::unordered_set<int32_t> ids;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '#') continue;
    split_line(line, tok);
    if (tok.size() < 2) return GU_ERR_PARSE;
    long long u, v;
    if (!to_int(tok[0], &u) || !to_int(tok[1], &v)) return GU_ERR_PARSE;
    if (u < 0 || v < 0) return GU_ERR_RANGE;
    if (Fv.size() >= (size_t)INT32_MAX - 1) return GU_ERR_CAPACITY;
    Fv.push_back((int32_t)u);
    Cv.push_back((int32_t)v);
    ids.insert((int32_t)u);
    ids.insert((int32_t)v);
  }
  const int64_t n = (int64_t)ids.size(), m = (int64_t)Fv.size();
  for (int64_t i = 0; i < m; i++)
    if (Fv[(size_t)i] >= n || Cv[(size_t)i] >= n) return GU_ERR_RANGE;  // ids not 0..n-1
  induce_R(n, Fv, m, Rv);
  *R = dup_out(Rv);
  *C = dup_out(Cv);
  if (F) *F = dup_out(Fv);
  *n_out = (int32_t)n;
  *m_out = m;
  return (*R && *C) ? GU_OK : GU_ERR_NOMEM;
}

// reverse (graph-utils/reverse.cpp:3-44): the transpose with its entries
// ordered by (target, source), R induced from the targets.  A counting sort
// by target over the entries in row order gives that order.
int gu_reverse(const int32_t *R, const int32_t *C, int32_t n, int64_t m, int32_t **R2, int32_t **C2,
               int32_t **F2) {
  if (!R || (!C && m > 0) || n < 0 || m < 0 || !R2 || !C2) return GU_ERR_INVALID;
  if (R[0] != 0 || (int64_t)R[n] != m) return GU_ERR_INVALID;
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  for (int64_t k = 0; k < m; k++) {
    if (C[k] < 0 || C[k] >= n) return GU_ERR_RANGE;
    cnt[(size_t)C[k] + 1]++;
  }
  for (int32_t v = 0; v < n; v++) cnt[(size_t)v + 1] += cnt[(size_t)v];
  std::vector<int32_t> Rv((size_t)n + 1), Cv((size_t)m), Fv((size_t)m);
  for (int32_t v = 0; v <= n; v++) Rv[(size_t)v] = (int32_t)cnt[(size_t)v];
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int32_t u = 0; u < n; u++)
    for (int32_t k = R[u]; k < R[u + 1]; k++) {
      const int64_t p = pos[(size_t)C[k]]++;
      Cv[(size_t)p] = u;
      Fv[(size_t)p] = C[k];
    }
  *R2 = dup_out(Rv);
  *C2 = dup_out(Cv);
  if (F2) *F2 = dup_out(Fv);
  return (*R2 && *C2) ? GU_OK : GU_ERR_NOMEM;
}

// host_graph::write_edgelist_to_file (parse.cpp:436-456): optional "n m"
// line, then "F[k] C[k]" per entry.
int gu_write_edgelist(const char *path, int32_t header, const int32_t *F, const int32_t *C, int32_t n,
                      int64_t m) {
  if (!path || n < 0 || m < 0 || (m > 0 && (!F || !C))) return GU_ERR_INVALID;
  FILE *f = std::fopen(path, "w");
  if (!f) return GU_ERR_IO;
  if (header) std::fprintf(f, "%d %lld\n", n, (long long)m);
  for (int64_t k = 0; k < m; k++) std::fprintf(f, "%d %d\n", F[k], C[k]);
  return std::fclose(f) == 0 ? GU_OK : GU_ERR_IO;
}

int gu_parse_metis(const char *path, int32_t **R, int32_t **C, int32_t **F,
                   int32_t *n_out, int64_t *m_out) {
  if (!path || !R || !C || !n_out || !m_out) return GU_ERR_INVALID;
  std::ifstream in(path, std::ifstream::in);
  if (!in.good()) return GU_ERR_IO;
  std::string line;
  bool first = true;
  int64_t n = 0, m = 0, node = 0, edge = 0;
  std::vector<int32_t> Rv, Cv, Fv;
  std::vector<std::string> tok;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '%') continue;
    split_line(line, tok);
    if (first) {
      long long a, b, c;
      if (tok.size() < 2 || !to_int(tok[0], &a) || !to_int(tok[1], &b)) return GU_ERR_PARSE;
      if (tok.size() > 3) return GU_ERR_WEIGHTED;
      if (tok.size() == 3) {
        if (!to_int(tok[2], &c)) return GU_ERR_PARSE;
        if (c != 0) return GU_ERR_WEIGHTED;
      }
      if (a < 0 || b < 0 || 2 * b >= INT32_MAX) return GU_ERR_PARSE;
      n = a;
      m = 2 * b;
      Rv.assign(n + 1, 0);
      Cv.assign(m, 0);
      Fv.assign(m, 0);
      first = false;
      node = 1;
    } else {
      if (node > n) {
        // Lines past the n-th adjacency line: blank ones are ignored (the
        // reference writes them past the end of R); any id there is an error.
        for (const std::string &s : tok)
          if (is_number(s)) return GU_ERR_PARSE;
        continue;
      }
      int64_t r = (int64_t)tok.size() + Rv[node - 1];
      for (const std::string &s : tok) {
        if (!is_number(s)) {
          r--;
          continue;
        }
        long long id;
        if (!to_int(s, &id)) return GU_ERR_PARSE;
        if (id < 1 || id > n) return GU_ERR_RANGE;
        if (edge >= m) return GU_ERR_PARSE;  // more entries than the header's m
        Cv[edge] = (int32_t)(id - 1);
        Fv[edge] = (int32_t)(node - 1);
        edge++;
      }
      Rv[node] = (int32_t)r;
      node++;
    }
  }
  if (first) return GU_ERR_PARSE;
  *R = dup_out(Rv);
  *C = dup_out(Cv);
  if (F) *F = dup_out(Fv);
  *n_out = (int32_t)n;
  *m_out = m;
  return GU_OK;
}

int gu_write_metis(const char *path, const int32_t *R, const int32_t *C, int32_t n,
                   int64_t m) {
  if (!path || !R || (!C && m > 0) || n < 0) return GU_ERR_INVALID;
  FILE *f = std::fopen(path, "w");
  if (!f) return GU_ERR_IO;
  std::fprintf(f, "%d %lld\n", n, (long long)(m / 2));
  std::string buf;
  for (int32_t v = 0; v < n; v++) {
    buf.clear();
    for (int32_t k = R[v]; k < R[v + 1]; k++) {
      if (k > R[v]) buf.push_back(' ');
      buf += std::to_string(C[k] + 1);
    }
    buf.push_back('\n');
    std::fwrite(buf.data(), 1, buf.size(), f);
  }
  return std::fclose(f) == 0 ? GU_OK : GU_ERR_IO;
}

int gu_gen_er(int32_t n, int64_t m_undirected, uint64_t seed, int32_t **R, int32_t **C,
              int32_t **F, int32_t *n_out, int64_t *m_out) {
  if (n < 0 || m_undirected < 0) return GU_ERR_INVALID;
  const int64_t N = (int64_t)n * (n - 1) / 2;
  if (m_undirected > N) return GU_ERR_INVALID;
  Xoshiro rng(seed);
  // Floyd's sampling of m distinct pair indices out of N.
  std::unordered_set<int64_t> chosen;
  std::vector<int64_t> order;
  chosen.reserve((size_t)m_undirected * 2 + 1);
  for (int64_t j = N - m_undirected; j < N; j++) {
    int64_t t = (int64_t)rng.below((uint64_t)j + 1);
    if (chosen.insert(t).second)
      order.push_back(t);
    else {
      chosen.insert(j);
      order.push_back(j);
    }
  }
  // Pair index -> (u, v), u < v, rows of the strict upper triangle.
  std::vector<int64_t> rowstart(n + 1, 0);
  for (int32_t u = 0; u < n; u++) rowstart[u + 1] = rowstart[u] + (n - 1 - u);
  std::vector<int32_t> eu, ev;
  eu.reserve(order.size());
  ev.reserve(order.size());
  for (int64_t idx : order) {
    int32_t u = (int32_t)(std::upper_bound(rowstart.begin(), rowstart.end(), idx) -
                          rowstart.begin() - 1);
    int32_t v = (int32_t)(u + 1 + (idx - rowstart[u]));
    eu.push_back(u);
    ev.push_back(v);
  }
  return build_csr(n, eu, ev, R, C, F, n_out, m_out);
}

int gu_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                uint64_t seed, int32_t **R, int32_t **C, int32_t **F, int32_t *n_out,
                int64_t *m_out) {
  if (scale < 1 || scale > 30 || edge_factor < 1) return GU_ERR_INVALID;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return GU_ERR_INVALID;
  const int32_t n = (int32_t)1 << scale;
  const int64_t E = (int64_t)edge_factor << scale;
  std::vector<int32_t> eu(E), ev(E);
  const int64_t block = 1 << 16;
  const int64_t nblocks = (E + block - 1) / block;
  const double ab = a + b, abc = a + b + c;
  parallel_for(nblocks, [&](int64_t lo, int64_t hi) {
    for (int64_t blk = lo; blk < hi; blk++) {
      Xoshiro rng(mix64(seed ^ mix64((uint64_t)blk + 1)));
      int64_t e1 = std::min(E, (blk + 1) * block);
      for (int64_t e = blk * block; e < e1; e++) {
        uint32_t u = 0, v = 0;
        for (int lvl = 0; lvl < scale; lvl++) {
          double r = rng.uniform();
          uint32_t bu = r >= ab ? 1u : 0u;              // quadrants c, d: lower half
          uint32_t bv = (r >= a && r < ab) || r >= abc;  // quadrants b, d: right half
          u = (u << 1) | bu;
          v = (v << 1) | bv;
        }
        eu[e] = (int32_t)u;
        ev[e] = (int32_t)v;
      }
    }
  });
  return build_csr(n, eu, ev, R, C, F, n_out, m_out);
}

int gu_gen_grid(int32_t side, double p_delete, uint64_t seed, int32_t lcc_only, int32_t **R,
                int32_t **C, int32_t **F, int32_t *n_out, int64_t *m_out) {
  if (side < 1 || side > 30000 || p_delete < 0 || p_delete > 1) return GU_ERR_INVALID;
  const int64_t nv = (int64_t)side * side;
  auto keep = [&](int64_t idx) {
    uint64_t x = mix64(seed ^ mix64((uint64_t)idx * 0x9E3779B97F4A7C15ull + 7));
    return (double)(x >> 11) * 0x1.0p-53 >= p_delete;
  };
  // Edge ids: horizontal (r,c)-(r,c+1) = r*(side-1)+c; vertical
  // (r,c)-(r+1,c) = side*(side-1) + r*side + c.
  const int64_t hcount = (int64_t)side * (side - 1);
  std::vector<int32_t> eu, ev;
  for (int64_t r = 0; r < side; r++)
    for (int64_t c2 = 0; c2 + 1 < side; c2++)
      if (keep(r * (side - 1) + c2)) {
        eu.push_back((int32_t)(r * side + c2));
        ev.push_back((int32_t)(r * side + c2 + 1));
      }
  for (int64_t r = 0; r + 1 < side; r++)
    for (int64_t c2 = 0; c2 < side; c2++)
      if (keep(hcount + r * side + c2)) {
        eu.push_back((int32_t)(r * side + c2));
        ev.push_back((int32_t)((r + 1) * side + c2));
      }
  int32_t *R0, *C0;
  int32_t n0;
  int64_t m0;
  int rc = build_csr((int32_t)nv, eu, ev, &R0, &C0, nullptr, &n0, &m0);
  if (rc != GU_OK) return rc;
  if (!lcc_only) {
    *R = R0;
    *C = C0;
    if (F) {
      std::vector<int32_t> Fv(m0);
      for (int32_t v = 0; v < n0; v++)
        for (int32_t k = R0[v]; k < R0[v + 1]; k++) Fv[k] = v;
      *F = dup_out(Fv);
    }
    *n_out = n0;
    *m_out = m0;
    return GU_OK;
  }
  // Components by BFS; keep the largest (ties: the one with the smallest id).
  std::vector<int32_t> comp(n0, -1), queue(n0);
  int32_t best = -1, ncomp = 0;
  int64_t best_size = -1;
  for (int32_t s = 0; s < n0; s++) {
    if (comp[s] >= 0) continue;
    int64_t h = 0, t = 0;
    queue[t++] = s;
    comp[s] = ncomp;
    while (h < t) {
      int32_t v = queue[h++];
      for (int32_t k = R0[v]; k < R0[v + 1]; k++)
        if (comp[C0[k]] < 0) {
          comp[C0[k]] = ncomp;
          queue[t++] = C0[k];
        }
    }
    if (t > best_size) {
      best_size = t;
      best = ncomp;
    }
    ncomp++;
  }
  std::vector<int32_t> newid(n0, -1);
  int32_t nn = 0;
  for (int32_t v = 0; v < n0; v++)
    if (comp[v] == best) newid[v] = nn++;
  std::vector<int32_t> Rv(nn + 1, 0), Cv, Fv;
  for (int32_t v = 0; v < n0; v++) {
    if (newid[v] < 0) continue;
    for (int32_t k = R0[v]; k < R0[v + 1]; k++) {
      Cv.push_back(newid[C0[k]]);  // relabelling is monotone: rows stay sorted
      Fv.push_back(newid[v]);
    }
    Rv[newid[v] + 1] = (int32_t)Cv.size();
  }
  std::free(R0);
  std::free(C0);
  *R = dup_out(Rv);
  *C = dup_out(Cv);
  if (F) *F = dup_out(Fv);
  *n_out = nn;
  *m_out = (int64_t)Cv.size();
  return GU_OK;
}

int gu_gen_zipf_counts(double s, int32_t cap, int64_t W, uint64_t seed, int32_t **counts,
                       int64_t *S) {
  if (s <= 0 || cap < 1 || W < 0 || W >= INT32_MAX || !counts || !S) return GU_ERR_INVALID;
  std::vector<double> cdf(cap);
  double acc = 0;
  for (int32_t k = 1; k <= cap; k++) {
    acc += std::pow((double)k, -s);
    cdf[k - 1] = acc;
  }
  for (auto &x : cdf) x /= acc;
  Xoshiro rng(seed);
  std::vector<int32_t> out;
  int64_t sum = 0;
  while (sum < W) {
    double u = rng.uniform();
    int32_t k = (int32_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    if (k > cap) k = cap;
    if (sum + k > W) k = (int32_t)(W - sum);
    out.push_back(k);
    sum += k;
  }
  *counts = dup_out(out);
  *S = (int64_t)out.size();
  return *counts ? GU_OK : GU_ERR_NOMEM;
}

int gu_sample_sources(const int32_t *R, int32_t n, int32_t k, uint64_t seed, int32_t *sources) {
  if (!R || !sources || n < 0 || k < 0) return GU_ERR_INVALID;
  std::vector<int32_t> cand;
  for (int32_t v = 0; v < n; v++)
    if (R[v + 1] > R[v]) cand.push_back(v);
  if (cand.empty()) return k == 0 ? GU_OK : GU_ERR_RANGE;
  Xoshiro rng(seed);
  for (int32_t i = 0; i < k; i++) sources[i] = cand[rng.below(cand.size())];
  return GU_OK;
}

}  // extern "C"
