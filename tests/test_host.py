"""CPU tests of the host side of the C ABI: symbol exports against
include/graph_utils_amd.h, struct layouts, the METIS reader (against the
reference's own parse_metis compiled into oracle/_ref where available), the
writer, the seeded generators and the [start,end) row mapping.  None of these
calls reaches a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

from conftest import FIXTURES, ROOT, graph_path

HEADER = os.path.join(ROOT, "include", "graph_utils_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(gu_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol(gu):
    declared = _declared_functions()
    assert len(declared) >= 28
    L = gu.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(gu.EXPORTS)


def test_status_strings(gu):
    codes = [0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11]
    msgs = [gu.lib().gu_strerror(c) for c in codes]
    assert all(isinstance(m, bytes) and m for m in msgs)
    assert len(set(msgs)) == len(msgs)
    assert gu.lib().gu_strerror(-1000) == b"unknown status"


def test_struct_layouts_match_header(gu, tmp_path):
    """ctypes mirrors of gu_config / gu_stats / gu_device_props have the C
    sizes and field offsets (compiled here with gcc against the header)."""
    classes = {"gu_config": gu.Config, "gu_stats": gu.Stats, "gu_device_props": gu.DeviceProps,
               "gu_server_config": gu.ServerConfig, "gu_query": gu.Query,
               "gu_query_result": gu.QueryResult, "gu_server_stats": gu.ServerStats,
               "gu_tune_result": gu.TuneResult}
    fields = {k: [f[0] for f in c._fields_] for k, c in classes.items()}
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "graph_utils_amd.h"',
           "int main(void) {"]
    for st, fl in fields.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (st, st))
        for f in fl:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, f, st, f))
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(c), "-o", str(exe)],
                   check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True,
                                                         capture_output=True, text=True)
               .stdout.splitlines())
    for st, cls in classes.items():
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f in fields[st]:
            assert int(out[st + "." + f]) == getattr(cls, f).offset, (st, f)


def test_calls_without_graph_are_rejected(gu):
    L = gu.lib()
    src = np.zeros(1, np.int32)
    assert L.gu_multi_search(None, src.ctypes.data, 1, None, None, None, None) == -1
    assert L.gu_betweenness(None, src.ctypes.data, 1, None, None, None, None) == -1
    assert L.gu_diameter(None, src.ctypes.data, 1, None, None, None) == -1
    assert L.gu_graph_destroy(None) == -1


# --------------------------------------------------------------------------
# METIS reader (parse.cpp:115-208)
# --------------------------------------------------------------------------
def _ref_or_skip(orc):
    if not orc.ref_parse_available():
        pytest.skip("oracle/_ref/libref_parse.so not built (needs the reference checkout)")


QUIRK_FILES = {
    # a leading blank makes the first id " 2" non-numeric: dropped (parse.cpp:143-154,188-193)
    "leading_blank": "3 2\n 2 3\n1\n1\n",
    # a doubled blank glues a blank onto the next id: dropped
    "double_blank": "3 2\n2  3\n1\n1\n",
    # CRLF: the last id of every line carries '\r': dropped
    "crlf": "3 2\r\n2 3\r\n1\r\n1\r\n",
    # tabs separate like blanks; comment lines are skipped
    "tabs_comments": "% c\n4 3\n2\t3\n% mid\n1\t4\n1\n2\n",
    # an empty adjacency line is a vertex without neighbours
    "empty_line": "4 1\n\n3\n2\n\n",
    # header with a zero weight flag is unweighted (parse.cpp:169)
    "flag_zero": "2 1 0\n2\n1\n",
}


@pytest.mark.parametrize("name", FIXTURES)
def test_parse_fixture_matches_reference_reader(gu, orc, name):
    _ref_or_skip(orc)
    hg = gu.HostGraph.parse_metis(graph_path(name))
    R, C, F, n, m, directed = orc.ref_parse_metis(graph_path(name))
    assert (hg.n, hg.m, hg.directed) == (n, m, directed)
    assert np.array_equal(hg.R, R) and np.array_equal(hg.C, C) and np.array_equal(hg.F, F)


@pytest.mark.parametrize("name", sorted(QUIRK_FILES))
def test_parse_quirks_match_reference_reader(gu, orc, name, tmp_path):
    _ref_or_skip(orc)
    p = tmp_path / (name + ".graph")
    p.write_bytes(QUIRK_FILES[name].encode())
    hg = gu.HostGraph.parse_metis(str(p))
    R, C, F, n, m, _ = orc.ref_parse_metis(str(p))
    assert (hg.n, hg.m) == (n, m)
    assert np.array_equal(hg.R, R) and np.array_equal(hg.C, C) and np.array_equal(hg.F, F)


def test_parse_quirks_expected_values(gu, tmp_path):
    """Known answers for the quirk files, independent of the reference build."""
    def parse(text):
        p = tmp_path / "q.graph"
        p.write_bytes(text.encode())
        return gu.HostGraph.parse_metis(str(p))

    g = parse(QUIRK_FILES["leading_blank"])
    assert g.R.tolist() == [0, 1, 2, 3] and g.C.tolist() == [2, 0, 0, 0]
    g = parse(QUIRK_FILES["crlf"])
    assert g.R.tolist() == [0, 1, 1, 1] and g.C.tolist()[:1] == [1]
    g = parse(QUIRK_FILES["empty_line"])
    assert g.R.tolist() == [0, 0, 1, 2, 2] and g.C.tolist() == [2, 1]
    g = parse(QUIRK_FILES["tabs_comments"])
    assert g.R.tolist() == [0, 2, 4, 5, 6] and g.C.tolist() == [1, 2, 0, 3, 0, 1]


@pytest.mark.parametrize("text,code", [
    ("3 2 1\n2 3\n1\n1\n", -6),        # weighted flag (parse.cpp:169-173 exits -2)
    ("3 2 0 1\n2 3\n1\n1\n", -6),      # four header fields (parse.cpp:164-168)
    ("x y\n", -5),                     # non-numeric header
    ("", -5),                          # no header at all
    ("3 1\n2 9\n1\n\n", -4),           # id beyond n
    ("2 1\n2\n1\n1\n", -5),            # ids on a line past the n-th
    ("2 1\n2 2 2\n1\n", -5),           # more entries than the header's m
])
def test_parse_errors_return_status(gu, tmp_path, text, code):
    p = tmp_path / "bad.graph"
    p.write_bytes(text.encode())
    with pytest.raises(gu.GuError) as e:
        gu.HostGraph.parse_metis(str(p))
    assert e.value.status == code


def test_parse_missing_file(gu, tmp_path):
    with pytest.raises(gu.GuError) as e:
        gu.HostGraph.parse_metis(str(tmp_path / "nope.graph"))
    assert e.value.status == -7


@pytest.mark.parametrize("name", FIXTURES)
def test_write_parse_round_trip(gu, name, tmp_path):
    hg = gu.HostGraph.parse_metis(graph_path(name))
    out = tmp_path / "rt.graph"
    hg.write_metis(str(out))
    text = out.read_bytes()
    assert b"\r" not in text and b"  " not in text and not re.search(rb"^ ", text, re.M)
    g2 = gu.HostGraph.parse_metis(str(out))
    assert np.array_equal(g2.R, hg.R) and np.array_equal(g2.C, hg.C)
    assert np.array_equal(g2.F, hg.F) and g2.m == hg.m


# --------------------------------------------------------------------------
# Seeded generators
# --------------------------------------------------------------------------
def _simple_symmetric(g):
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.R))
    assert np.all(g.C >= 0) and np.all(g.C < n)
    assert not np.any(rows == g.C), "self-loop"
    A = sp.csr_matrix((np.ones(len(g.C)), (rows, g.C)), shape=(n, n))
    assert A.nnz == len(g.C), "duplicate edge"
    assert (A != A.T).nnz == 0, "not symmetric"
    assert np.array_equal(g.F, rows)
    for v in range(0, n, max(1, n // 97)):
        row = g.C[g.R[v]:g.R[v + 1]]
        assert np.all(np.diff(row) > 0), "row not sorted"
    return A


def test_er_matches_committed_fixture(gu):
    """C1's graph: ER(198, 2742), seed 0, as committed with the golden vectors."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "er198.npz"))
    g = gu.HostGraph.er(198, 2742, seed=0)
    assert g.n == 198 and g.m == 5484
    assert np.array_equal(g.R, z["R"]) and np.array_equal(g.C, z["C"])
    _simple_symmetric(g)


def test_er_seeds(gu):
    a = gu.HostGraph.er(500, 4000, seed=1)
    b = gu.HostGraph.er(500, 4000, seed=1)
    c = gu.HostGraph.er(500, 4000, seed=2)
    assert np.array_equal(a.C, b.C) and not np.array_equal(a.C, c.C)
    _simple_symmetric(a)
    full = gu.HostGraph.er(20, 190, seed=3)          # complete graph K20
    assert np.array_equal(np.diff(full.R), np.full(20, 19))
    with pytest.raises(gu.GuError):
        gu.HostGraph.er(20, 191, seed=3)


def test_rmat_properties(gu):
    a = gu.HostGraph.rmat(12, 16, seed=5)
    b = gu.HostGraph.rmat(12, 16, seed=5)
    c = gu.HostGraph.rmat(12, 16, seed=6)
    assert a.n == 4096
    assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C)
    assert not np.array_equal(a.C, c.C)
    _simple_symmetric(a)
    deg = np.diff(a.R)
    # skewed: vertex 0 gets the a-quadrant mass, isolated vertices exist
    assert deg[0] == deg.max() and np.count_nonzero(deg == 0) > 0
    assert a.m <= 2 * 16 * 4096


def test_grid_intact_and_lcc(gu):
    g = gu.HostGraph.grid(16, 0.0, seed=0)
    assert g.n == 256 and g.m == 2 * 2 * 16 * 15
    _simple_symmetric(g)
    deg = np.diff(g.R)
    assert deg.min() == 2 and deg.max() == 4
    h = gu.HostGraph.grid(64, 0.10, seed=4)
    A = _simple_symmetric(h)
    assert connected_components(A, directed=False)[0] == 1
    assert h.n <= 64 * 64 and h.m < 2 * 2 * 64 * 63
    sparse = gu.HostGraph.grid(64, 0.45, seed=4)      # near the percolation point
    assert connected_components(_simple_symmetric(sparse), directed=False)[0] == 1
    assert sparse.n < 64 * 64
    full = gu.HostGraph.grid(64, 0.45, seed=4, lcc_only=False)
    assert full.n == 64 * 64
    assert connected_components(_simple_symmetric(full), directed=False)[0] > 1
    assert np.array_equal(gu.HostGraph.grid(64, 0.10, seed=4).C, h.C)


def test_zipf_counts(gu):
    W = 1 << 16
    c = gu.zipf_counts(1.1, 1 << 12, W, seed=7)
    assert c.dtype == np.int32 and int(c.astype(np.int64).sum()) == W
    assert c.min() >= 1 and c.max() <= (1 << 12)
    assert np.array_equal(c, gu.zipf_counts(1.1, 1 << 12, W, seed=7))
    # heavy head: count 1 is the most frequent value
    vals, cnt = np.unique(c, return_counts=True)
    assert vals[np.argmax(cnt)] == 1


def test_sample_sources(gu):
    g = gu.HostGraph.rmat(12, 16, seed=5)
    s = g.sample_sources(256, seed=9)
    assert np.all(np.diff(g.R)[s] > 0)
    assert np.array_equal(s, g.sample_sources(256, seed=9))
    assert not np.array_equal(s, g.sample_sources(256, seed=10))


@pytest.mark.parametrize("start,end,G", [(0, 5, 256), (0, 198, 256), (0, 300, 256),
                                         (10, 700, 256), (3, 4, 1), (0, 1000, 7),
                                         (5, 5, 256)])
def test_row_sources_match_oracle(gu, orc, start, end, G):
    assert np.array_equal(gu.row_sources(start, end, G), orc.rows_for_range(start, end, G))


# --------------------------------------------------------------------------
# Matrix Market and SNAP readers, edge-list writer (parse.cpp:210-456)
# --------------------------------------------------------------------------
def _mm_text(hg, banner=True, tabs=False):
    """Lower-triangle pattern Matrix Market text of an undirected graph."""
    rows = [(int(v) + 1, int(u) + 1) for u in range(hg.n) for v in hg.C[hg.R[u]:hg.R[u + 1]] if v > u]
    sep = "\t" if tabs else " "
    head = "%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n" if banner else ""
    return head + "%d%s%d%s%d\n" % (hg.n, sep, hg.n, sep, len(rows)) + "".join(
        "%d%s%d\n" % (a, sep, b) for a, b in rows)


@pytest.mark.parametrize("name", FIXTURES)
def test_market_matches_reference_reader(gu, orc, name, tmp_path):
    _ref_or_skip(orc)
    hg = gu.HostGraph.parse_metis(graph_path(name))
    p = tmp_path / (name + ".mtx")
    p.write_text(_mm_text(hg, tabs=(name == "test3")))
    g = gu.HostGraph.parse_market(str(p))
    R, C, F, n, m = orc.ref_parse_market(str(p))
    assert (g.n, g.m) == (n, m)
    assert np.array_equal(g.R, R) and np.array_equal(g.C, C) and np.array_equal(g.F, F)
    # symmetrised and sorted: the same CSR as the METIS file's (sorted rows)
    assert np.array_equal(g.R, hg.R)
    assert np.array_equal(g.C, np.concatenate([np.sort(hg.C[hg.R[u]:hg.R[u + 1]]) for u in range(hg.n)]))


def test_market_quirks_match_reference(gu, orc, tmp_path):
    _ref_or_skip(orc)
    cases = {
        "selfloop_and_dups": "4 4 4\n1 1\n2 1\n2 1\n4 3\n",        # kept twice each
        "non_square": "3 5 2\n1 2\n3 2\n",                         # the reference only warns
        "isolated_tail": "6 6 2\n2 1\n3 2\n",                      # rows 3..5 empty
        "more_than_header": "3 3 1\n2 1\n3 2\n",                   # first m entries after sorting
    }
    for name, text in cases.items():
        p = tmp_path / (name + ".mtx")
        p.write_text(text)
        g = gu.HostGraph.parse_market(str(p))
        R, C, F, n, m = orc.ref_parse_market(str(p))
        assert (g.n, g.m) == (n, m), name
        assert np.array_equal(g.R, R) and np.array_equal(g.C, C) and np.array_equal(g.F, F), name


@pytest.mark.parametrize("text,code", [
    ("3 3\n1 2\n", -5),              # header needs three fields (parse.cpp exits -2)
    ("3 3 1\n1 2 1\n", -5),          # values: only pattern matrices
    ("3 3 2\n1 2\n", -5),            # fewer entries than the header
    ("3 3 1\n1 4\n", -4),            # id beyond n
])
def test_market_errors(gu, tmp_path, text, code):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(gu.GuError) as e:
        gu.HostGraph.parse_market(str(p))
    assert e.value.status == code


SNAP_FILES = {
    "sorted_tabs": "# c\n0\t2\n0\t3\n1\t3\n2\t0\n3\t1\n",
    "extra_fields": "0 1 7\n1 2 8\n2 0 9\n",
    "unsorted": "2 0\n0 1\n1 2\n0 2\n",       # the reference's R for an unsorted file
    "selfloop_crlf": "0 0\r\n0 1\r\n1 0\r\n",
    "isolated_target": "0 1\n0 2\n",
}


@pytest.mark.parametrize("name", sorted(SNAP_FILES))
@pytest.mark.parametrize("header", [False, True])
def test_snap_matches_reference_reader(gu, orc, name, header, tmp_path):
    _ref_or_skip(orc)
    p = tmp_path / (name + ".txt")
    p.write_bytes(SNAP_FILES[name].encode())
    g = gu.HostGraph.parse_snap(str(p), header)
    R, C, F, n, m, directed = orc.ref_parse_snap(str(p), header)
    assert directed and g.directed
    assert (g.n, g.m) == (n, m)
    assert np.array_equal(g.R, R) and np.array_equal(g.C, C) and np.array_equal(g.F, F)


def test_snap_reference_fixture(gu, orc):
    """graphs/test7.txt of the reference: 11 vertices, 13 directed entries."""
    path = os.path.join(ROOT, "tests", "golden", "graphs", "test7.txt")
    g = gu.HostGraph.parse_snap(path)
    assert (g.n, g.m) == (11, 13)
    assert g.R.tolist() == [0, 3, 5, 8, 9, 10, 11, 11, 13, 13, 13, 13]
    assert g.C[:3].tolist() == [2, 3, 4]
    if orc.ref_parse_available():
        R, C, F, n, m, _ = orc.ref_parse_snap(path)
        assert np.array_equal(g.R, R) and np.array_equal(g.C, C)


@pytest.mark.parametrize("text,code", [
    ("0\n", -5),                 # one field
    ("0 1\n5 0\n", -4),          # id 5 with only 3 distinct ids
    ("a b\n", -5),
    ("0 -1\n", -4),
])
def test_snap_errors(gu, tmp_path, text, code):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(gu.GuError) as e:
        gu.HostGraph.parse_snap(str(p))
    assert e.value.status == code


@pytest.mark.parametrize("header", [False, True])
def test_edgelist_round_trip(gu, orc, tmp_path, header):
    hg = gu.HostGraph.parse_metis(graph_path("test3"))
    p = tmp_path / "el.txt"
    hg.write_edgelist(str(p), header)
    lines = p.read_text().splitlines()
    assert len(lines) == hg.m + (1 if header else 0)
    if header:
        assert lines[0] == "%d %d" % (hg.n, hg.m)
        # the reference's reader ignores its header flag: "n m" becomes an
        # entry whose ids lie outside the graph (the reference then indexes
        # past R); this reader refuses it
        with pytest.raises(gu.GuError) as e:
            gu.HostGraph.parse_snap(str(p), True)
        assert e.value.status == -4
        return
    g = gu.HostGraph.parse_snap(str(p))   # rows ascend: the same CSR back
    assert np.array_equal(g.R, hg.R) and np.array_equal(g.C, hg.C)
    if orc.ref_parse_available():
        R, C, F, n, m, _ = orc.ref_parse_snap(str(p))
        assert np.array_equal(g.R, R) and np.array_equal(g.C, C)


# --------------------------------------------------------------------------
# reverse and DAG reachability (graph-utils/reverse.cpp, dag_reachability.cpp)
# --------------------------------------------------------------------------
def _random_dag(n, m, seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    a, b = np.minimum(u, v)[keep], np.maximum(u, v)[keep]
    perm = rng.permutation(n)          # hide the topological order in the labels
    a, b = perm[a], perm[b]
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    R = np.zeros(n + 1, dtype=np.int32)
    np.add.at(R, a + 1, 1)
    R = np.cumsum(R).astype(np.int32)
    return R, b.astype(np.int32), a.astype(np.int32)


@pytest.mark.parametrize("case", ["jazz", "test3", "dag"])
def test_reverse_matches_reference(gu, orc, case):
    _ref_or_skip(orc)
    if case == "dag":
        R, C, F = _random_dag(300, 1500, 3)
        hg = gu.HostGraph(R, C, F, directed=True)
    else:
        hg = gu.HostGraph.parse_metis(graph_path(case))
    r = hg.reverse()
    R2, C2, F2 = orc.ref_reverse(hg.R, hg.C, hg.F)
    assert np.array_equal(r.R, R2) and np.array_equal(r.C, C2) and np.array_equal(r.F, F2)
    rr = r.reverse()                  # reversing twice gives rows in ascending order
    assert np.array_equal(rr.R, hg.R)


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_reachability_matches_reference_dag(orc, seed):
    """The oracle's BFS reachability (finite distance, not the source) equals
    the reference's recursive DAG reachability: the check the GPU form of
    get_reachability_dag is held to."""
    _ref_or_skip(orc)
    R, C, F = _random_dag(200, 700, seed)
    ref = orc.ref_reach_dag(R, C, F)
    for s in range(200):
        d = orc.bfs(R, C, s)
        got = np.nonzero((d != orc.INT_MAX) & (d > 0))[0]
        assert np.array_equal(got, ref[s]), s


def test_missing_library_fails_loudly():
    """Without the HIP library the package raises instead of computing
    anything another way (no CPU fallback exists)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import __graft_entry__\n"
        "pkg = __graft_entry__.load_package()\n"
        "pkg.LIB_PATH = '/nonexistent/libgraph_utils_amd.so'\n"
        "pkg._lib = None\n"
        "try:\n"
        "    pkg.HostGraph.er(10, 5)\n"
        "except ImportError as e:\n"
        "    print('ImportError:', e)\n"
        "    sys.exit(0)\n"
        "sys.exit(1)\n" % ROOT)
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "not built" in r.stdout


def _create(gu, R, C, directed):
    """gu_graph_create's status; a handle made on a GPU machine is destroyed."""
    R = np.ascontiguousarray(R, dtype=np.int32)
    C = np.ascontiguousarray(C, dtype=np.int32)
    h = ctypes.c_void_p()
    rc = gu.lib().gu_graph_create(R.ctypes.data_as(ctypes.c_void_p), C.ctypes.data_as(ctypes.c_void_p), None,
                                  len(R) - 1, len(C), directed, ctypes.byref(h))
    if rc == 0:
        gu.lib().gu_graph_destroy(h)
    return rc


def test_graph_create_refuses_asymmetric_undirected(gu):
    """directed = 0 promises a symmetric CSR (the pulls read the adjacency as
    in-edges); one that is not is refused on the host before any device
    call, so this holds with or without a GPU."""
    R = np.array([0, 2, 3, 4], np.int32)   # 0->1, 0->2, 1->0, 2->1: 2->1 has no 1->2
    C = np.array([1, 2, 0, 1], np.int32)
    assert _create(gu, R, C, 0) == gu.ERR_INVALID
    ok = (gu.OK, gu.ERR_NODEVICE)
    assert _create(gu, R, C, 1) in ok                    # the same CSR as a directed graph
    Rs = np.array([0, 2, 4, 6], np.int32)                # the symmetric triangle
    Cs = np.array([1, 2, 0, 2, 0, 1], np.int32)
    assert _create(gu, Rs, Cs, 0) in ok
    # parallel edges must come in pairs too: 0-1 twice, 1-0 once
    Rp = np.array([0, 2, 3], np.int32)
    Cp = np.array([1, 1, 0], np.int32)
    assert _create(gu, Rp, Cp, 0) == gu.ERR_INVALID
    assert _create(gu, np.array([0, 2, 4], np.int32), np.array([1, 1, 0, 0], np.int32), 0) in ok


def test_over_long_market_file_is_not_a_symmetric_graph(gu, tmp_path):
    """parse_market keeps the first m sorted entries of a file with more
    entries than its header (as parse.cpp does), which drops mirror entries;
    the resulting CSR cannot be created as an undirected graph."""
    p = tmp_path / "long.mtx"
    p.write_text("3 3 1\n3 1\n2 1\n")  # sorted entries (0,1) (0,2) (1,0) (2,0); m = 2 keeps the first two
    g = gu.HostGraph.parse_market(str(p))
    assert g.m == 2 and list(g.C) == [1, 2]
    assert _create(gu, g.R, g.C, 0) == gu.ERR_INVALID
    # every fixture and generator output is symmetric
    for hg in (gu.HostGraph.parse_metis(graph_path("jazz")), gu.HostGraph.rmat(10, 8, seed=3),
               gu.HostGraph.grid(20, 0.1, seed=4)):
        assert _create(gu, hg.R, hg.C, 0) in (gu.OK, gu.ERR_NODEVICE)
