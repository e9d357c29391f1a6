"""Pins the CPU oracle (oracle/oracle.c) before anything is compared with it.

Each restated function is checked against something that does not share its
code: closed forms (star, path, tree, intact grid binomials mod 2^64), brute
force path enumeration, scipy's BFS, networkx's Brandes, numpy's
searchsorted and a simulation of the reference's block-per-source loop.
The committed golden vectors (tests/golden/*.npz) are checked last, so a
change to the oracle that still passes the independent checks cannot drift
from the fixtures either.
"""
import itertools
import os
from math import comb

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

from conftest import FIXTURES, GOLDEN, graph_path

INT_MAX = 2147483647


def load_ref_graph(orc, name):
    """Fixture graph via the reference's own parse_metis when it is built,
    else via a minimal independent reader."""
    path = graph_path(name)
    if orc.ref_parse_available():
        R, C, F, n, m, _ = orc.ref_parse_metis(path)
        return R, C
    return simple_metis(path)


def simple_metis(path):
    rows = []
    header = None
    with open(path) as f:
        for line in f:
            if line.startswith("%"):
                continue
            toks = line.split()
            if header is None:
                header = toks
                continue
            rows.append([int(t) - 1 for t in toks])
    n = int(header[0])
    rows += [[] for _ in range(n - len(rows))]
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(r) for r in rows[:n]])
    C = np.array([c for r in rows[:n] for c in r], dtype=np.int32)
    return R, C


def to_nx(R, C):
    g = nx.Graph()
    n = len(R) - 1
    g.add_nodes_from(range(n))
    for v in range(n):
        for k in range(R[v], R[v + 1]):
            g.add_edge(v, int(C[k]))
    return g


def csr_from_edges(n, edges):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(a) for a in adj])
    C = np.array([c for a in adj for c in sorted(a)], dtype=np.int32)
    return R, C


def grid_graph(W):
    edges = []
    for r in range(W):
        for c in range(W):
            v = r * W + c
            if c + 1 < W:
                edges.append((v, v + 1))
            if r + 1 < W:
                edges.append((v, v + W))
    return csr_from_edges(W * W, edges)


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    return csr_from_edges(n, edges)


def brute_force_paths(R, C, s):
    """Shortest-path distances and counts by enumerating simple paths."""
    n = len(R) - 1
    best = [INT_MAX] * n
    cnt = [0] * n
    best[s], cnt[s] = 0, 1

    def dfs(v, depth, seen):
        for k in range(R[v], R[v + 1]):
            w = int(C[k])
            if w in seen:
                continue
            if depth + 1 < best[w]:
                best[w], cnt[w] = depth + 1, 1
            elif depth + 1 == best[w]:
                cnt[w] += 1
            seen.add(w)
            dfs(w, depth + 1, seen)
            seen.remove(w)

    dfs(s, 0, {s})
    return np.array(best, dtype=np.int64), np.array(cnt, dtype=np.uint64)


# ---------------------------------------------------------------- distances


@pytest.mark.parametrize("name", FIXTURES)
def test_bfs_matches_scipy(orc, name):
    R, C = load_ref_graph(orc, name)
    n = len(R) - 1
    A = sp.csr_matrix((np.ones(len(C)), C, R), shape=(n, n))
    D = shortest_path(A, unweighted=True, directed=True)
    for s in range(n):
        d = orc.bfs(R, C, s)
        exp = np.where(np.isinf(D[s]), INT_MAX, D[s]).astype(np.int64)
        assert np.array_equal(d.astype(np.int64), exp)


def test_bfs_closed_forms(orc):
    R, C = load_ref_graph(orc, "test5")  # path P7
    for s in range(7):
        assert np.array_equal(orc.bfs(R, C, s), np.abs(np.arange(7) - s))
    R, C = load_ref_graph(orc, "test6")  # 6-star plus 3 isolated vertices
    d = orc.bfs(R, C, 0)
    assert list(d) == [0, 1, 1, 1, 1, 1, INT_MAX, INT_MAX, INT_MAX]
    d = orc.bfs(R, C, 7)  # degree-0 source: its row stays 0 / INT_MAX
    assert d[7] == 0 and np.all(np.delete(d, 7) == INT_MAX)


# -------------------------------------------------------------- path counts


def test_sigma_grid_binomials_mod_2_64(orc):
    W = 40  # counts reach C(78, 39) ~ 2.6e22 > 2^64: pins the wrap
    R, C = grid_graph(W)
    d, s = orc.sigma(R, C, 0)
    for r in range(W):
        for c in range(W):
            assert d[r * W + c] == r + c
            assert int(s[r * W + c]) == comb(r + c, r) % (1 << 64)


@pytest.mark.parametrize("seed", range(6))
def test_sigma_brute_force(orc, seed):
    R, C = random_graph(10, 0.35, seed)
    for s in range(10):
        d, sig = orc.sigma(R, C, s)
        bd, bc = brute_force_paths(R, C, s)
        assert np.array_equal(d.astype(np.int64), bd)
        assert np.array_equal(sig, bc)


@pytest.mark.parametrize("name", ["test1", "test2", "test6"])
def test_sigma_trees_are_one(orc, name):
    R, C = load_ref_graph(orc, name)
    for s in range(len(R) - 1):
        d, sig = orc.sigma(R, C, s)
        assert np.all(sig[d != INT_MAX] == 1) and np.all(sig[d == INT_MAX] == 0)


# ------------------------------------------------------------- dependencies


CLOSED_BC = {
    "test1": [38, 24, 0, 36, 14, 0, 0, 0, 0],
    "test3": [0, 7, 2, 1, 2],
    "test4": [6, 0, 6, 30, 12, 12, 14, 0, 0],
    "test5": [0, 10, 16, 18, 16, 10, 0],
}


@pytest.mark.parametrize("name", sorted(CLOSED_BC))
def test_bc_closed_forms(orc, name):
    R, C = load_ref_graph(orc, name)
    n = len(R) - 1
    bc = orc.bc_f64(R, C, np.arange(n))
    assert np.allclose(bc, CLOSED_BC[name], rtol=0, atol=1e-9)


def test_bc_star_and_tree_formula(orc):
    R, C = load_ref_graph(orc, "test2")
    bc = orc.bc_f64(R, C, np.arange(6))
    assert bc[0] == 20 and np.all(bc[1:] == 0)  # (n-1)(n-2)
    R, C = load_ref_graph(orc, "test1")  # tree: (n-1)^2 - sum |component of T-v|^2
    g = to_nx(R, C)
    bc = orc.bc_f64(R, C, np.arange(9))
    for v in range(9):
        h = g.copy()
        h.remove_node(v)
        comp = [len(c) for c in nx.connected_components(h)]
        assert bc[v] == pytest.approx(64 - sum(x * x for x in comp))


@pytest.mark.parametrize("name", FIXTURES)
def test_bc_matches_networkx(orc, name):
    R, C = load_ref_graph(orc, name)
    n = len(R) - 1
    ref = nx.betweenness_centrality(to_nx(R, C), normalized=False)
    bc = orc.bc_f64(R, C, np.arange(n))
    assert np.allclose(bc, [2 * ref[v] for v in range(n)], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("name", ["jazz", "test4"])
def test_brandes_f32_close_to_f64(orc, name):
    R, C = load_ref_graph(orc, name)
    for s in range(0, len(R) - 1, 5):
        d32 = orc.brandes_f32(R, C, s).astype(np.float64)
        d64 = orc.brandes_f64(R, C, s)
        assert d32[s] == 0 and d64[s] == 0
        assert np.allclose(d32, d64, rtol=1e-5, atol=1e-6)


def test_sampled_bc_is_sum_of_rows(orc):
    R, C = load_ref_graph(orc, "jazz")
    src = [3, 50, 120, 197]
    bc = orc.bc_f64(R, C, src)
    assert np.allclose(bc, sum(orc.brandes_f64(R, C, s) for s in src), rtol=0, atol=1e-9)


# ---------------------------------------------------------------- diameter


def test_eccentricity(orc):
    R, C = load_ref_graph(orc, "test5")
    assert max(orc.eccentricity(R, C, s) for s in range(7)) == 6
    R, C = load_ref_graph(orc, "test2")
    assert max(orc.eccentricity(R, C, s) for s in range(6)) == 2
    R, C = load_ref_graph(orc, "jazz")
    assert max(orc.eccentricity(R, C, s) for s in range(198)) == 6
    R, C = grid_graph(9)
    assert orc.eccentricity(R, C, 0) == 16
    assert max(orc.eccentricity(R, C, s) for s in range(81)) == 16


# --------------------------------------------------------------------- LBS


LBS_CASES = [
    [0, 3, 0, 5, 1, 0, 2],
    [0, 0, 4, 0, 0, 0, 7, 1, 0, 0],
    [5],
    [0, 0, 0],
    [1] * 100,
]


@pytest.mark.parametrize("counts", LBS_CASES)
def test_lbs_against_searchsorted(orc, counts):
    scan, W = orc.exclusive_scan(counts)
    assert W == sum(counts)
    assert np.array_equal(scan, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32))
    res = orc.lbs(W, scan)
    assert np.array_equal(res, np.searchsorted(scan, np.arange(W), side="right") - 1)
    # every item belongs to a non-empty segment and ranks are in range
    for i, seg in enumerate(res):
        assert counts[seg] > 0 and 0 <= i - scan[seg] < counts[seg]


def test_lbs_random(orc):
    rng = np.random.default_rng(7)
    c = rng.integers(0, 50, 3000)
    c[rng.random(3000) < 0.2] = 0
    scan, W = orc.exclusive_scan(c)
    res = orc.lbs(W, scan)
    assert np.array_equal(res, np.searchsorted(scan, np.arange(W), side="right") - 1)


# ---------------------------------------------------------------- row map


def simulate_rows(start, end, G):
    """The reference's kernels: block b runs sources start+b, start+b+G, ...
    and its row keeps the last one (shuffle_based.cuh:104)."""
    rows = min(end - start, G)
    out = [None] * rows
    for b in range(rows):
        for s in range(start + b, end, G):
            out[b] = s
    return out


@pytest.mark.parametrize("start,end,G", [(0, 9, 256), (0, 198, 256), (0, 1000, 256), (5, 700, 64),
                                         (3, 4, 1), (0, 256, 256), (0, 257, 256)])
def test_rows_for_range(orc, start, end, G):
    assert list(orc.rows_for_range(start, end, G)) == simulate_rows(start, end, G)


# ---------------------------------------------------------- golden vectors


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
def test_golden_vectors(orc, name):
    path = os.path.join(GOLDEN, name + ".npz")
    z = np.load(path, allow_pickle=False)
    R, C = z["R"], z["C"]
    n = len(R) - 1
    for s in range(n):
        d, sig = orc.sigma(R, C, s)
        assert np.array_equal(d, z["dist"][s])
        assert np.array_equal(sig, z["sigma"][s])
        assert np.array_equal(orc.brandes_f32(R, C, s), z["delta_f32"][s])
    assert np.allclose(orc.bc_f64(R, C, np.arange(n)), z["bc_f64"], rtol=1e-12, atol=1e-9)


def test_planted_mistake_is_caught(orc):
    """A Brandes restatement that forgets the v != source skip (bc.cu:50)
    would credit the source; the star closed form catches it."""
    R, C = load_ref_graph(orc, "test2")
    wrong = np.zeros(6)
    for s in range(6):
        d = orc.brandes_f64(R, C, s)
        d = d.copy()
        d[s] += sum(1 for v in range(6) if v != s)  # what the unskipped form adds
        wrong += d
    assert not np.allclose(wrong, [20, 0, 0, 0, 0, 0])


def test_bc_f64_multithreaded_matches_serial(orc):
    R, C = load_ref_graph(orc, "jazz")
    src = np.arange(0, 198, 3)
    assert np.allclose(orc.bc_f64_mt(R, C, src, 4), orc.bc_f64(R, C, src), rtol=1e-13, atol=1e-10)


# ---- the full-size row checker (or_check_search_rows) and the threaded rows


def _er_rows(orc, n=300, p=0.012, seed=5, k=12):
    R, C = random_graph(n, p, seed)
    src = np.random.default_rng(seed).integers(0, n, k).astype(np.int32)
    d, s, e = orc.rows_mt(R, C, src, sigma=True, ecc=True, nthreads=3)
    return R, C, src, d, s, e


def test_rows_mt_matches_serial_oracle(orc):
    R, C, src, d, s, e = _er_rows(orc)
    for i, x in enumerate(src):
        od, os_ = orc.sigma(R, C, int(x))
        assert np.array_equal(d[i], od) and np.array_equal(s[i], os_)
        assert e[i] == orc.eccentricity(R, C, int(x))
    assert np.array_equal(orc.eccentricities_mt(R, C, src, nthreads=4), e)


def test_check_rows_accepts_oracle_and_scipy_rows(orc):
    R, C, src, d, s, _ = _er_rows(orc)
    assert orc.check_search_rows(R, C, src, d, s, nthreads=2) == (0, None)
    # distances from an independent BFS (scipy), distances-only check
    n = len(R) - 1
    A = sp.csr_matrix((np.ones(len(C)), C, R), shape=(n, n))
    sd = shortest_path(A, unweighted=True, indices=src)
    sd = np.where(np.isinf(sd), INT_MAX, sd).astype(np.int32)
    assert orc.check_search_rows(R, C, src, sd, None) == (0, None)
    # the intact grid's closed forms: Manhattan distance and binomials
    W = 30
    Rg, Cg = grid_graph(W)
    r, c = np.divmod(np.arange(W * W), W)
    gd = (r + c).astype(np.int32)[None, :]
    gs = np.array([comb(int(a + b), int(a)) % (1 << 64) for a, b in zip(r, c)], dtype=np.uint64)[None, :]
    assert orc.check_search_rows(Rg, Cg, np.array([0], np.int32), gd, gs) == (0, None)


@pytest.mark.parametrize("mistake", ["d_plus_one", "d_minus_one", "unreached_finite", "reached_inf",
                                     "sigma_off", "src_dist", "src_sigma", "unreached_sigma",
                                     "swap_rows", "level_shift"])
def test_check_rows_catches_planted_mistakes(orc, mistake):
    """Each plausible kernel mistake, planted in oracle rows on the CPU, must
    be reported, with its location."""
    R, C, src, d, s, _ = _er_rows(orc)
    d, s = d.copy(), s.copy()
    row = 3
    reach = np.flatnonzero((d[row] != INT_MAX) & (np.arange(len(R) - 1) != src[row]))
    unreach = np.flatnonzero(d[row] == INT_MAX)
    v = int(reach[len(reach) // 2])
    if mistake == "d_plus_one":
        d[row, v] += 1
    elif mistake == "d_minus_one":
        d[row, v] -= 1
    elif mistake == "unreached_finite":
        assert len(unreach) > 0
        d[row, unreach[0]] = 5
    elif mistake == "reached_inf":
        d[row, v] = INT_MAX
    elif mistake == "sigma_off":
        s[row, v] += np.uint64(1)
    elif mistake == "src_dist":
        d[row, src[row]] = 1
    elif mistake == "src_sigma":
        s[row, src[row]] = np.uint64(2)
    elif mistake == "unreached_sigma":
        assert len(unreach) > 0
        s[row, unreach[0]] = np.uint64(1)
    elif mistake == "swap_rows":  # rows stored under the wrong source
        d[[row, row + 1]] = d[[row + 1, row]]
        s[[row, row + 1]] = s[[row + 1, row]]
    elif mistake == "level_shift":  # every distance one too large
        fin = d[row] != INT_MAX
        d[row, fin] += 1
        d[row, src[row]] = 0
    viol, where = orc.check_search_rows(R, C, src, d, s)
    assert viol > 0
    assert where[0] in (row, row + 1)


def test_teps_units_is_component_degree_sum(orc):
    """or_teps_units (the GTEPS numerator) against an independent count: the
    degree sum of the source's connected component from scipy."""
    from scipy.sparse.csgraph import connected_components

    for seed in (1, 2, 3):
        R, C = random_graph(400, 0.004, seed)
        n = len(R) - 1
        A = sp.csr_matrix((np.ones(len(C)), C, R), shape=(n, n))
        _, lab = connected_components(A, directed=False)
        deg = np.diff(R).astype(np.int64)
        comp_deg = np.bincount(lab, weights=deg).astype(np.int64)
        for s in np.random.default_rng(seed).integers(0, n, 20):
            assert orc.teps_units(R, orc.bfs(R, C, int(s))) == comp_deg[lab[s]]


@pytest.mark.parametrize("W,wraps", [(34, False), (35, True)])
def test_sigma_wraps_grid_threshold(orc, W, wraps):
    """From a corner of the intact W x W grid the largest path count is
    C(2W-2, W-1) at the far corner: it fits in 64 bits for W = 34
    (7.2e18) and not for W = 35 (2.8e19)."""
    assert (comb(2 * W - 2, W - 1) >= 1 << 64) == wraps
    R, C = grid_graph(W)
    assert orc.sigma_wraps(R, C, 0) == wraps
    # from the centre every count is far below 2^64 at W = 34
    if not wraps:
        assert not orc.sigma_wraps(R, C, (W // 2) * W + W // 2)


def test_sigma_wraps_fixtures_never(orc):
    for name in FIXTURES:
        R, C = load_ref_graph(orc, name)
        assert not any(orc.sigma_wraps(R, C, s) for s in range(len(R) - 1))
