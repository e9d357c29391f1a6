"""GPU parity: the HIP path against the CPU oracle, through the C ABI.

Distances, path counts and LBS outputs must match bit for bit; dependencies
and BC must be within 1e-5 relative of the FP64 Brandes oracle (BC exactly 0
where the oracle's is 0).
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import FIXTURES, announce, graph_path

pytestmark = pytest.mark.gpu

VARIANTS = ["linear", "edge_parallel", "warp", "scan", "cta_warp", "shuffle"]
DIRS = {"auto": 0, "topdown": 1, "bottomup": 2}
BC_RTOL = 1e-5


def _er_graph(gu):
    hg = gu.HostGraph.er(198, 2742, seed=0)
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "er198.graph")
        hg.write_metis(p)
        return gu.HostGraph.parse_metis(p)


def _load(gu, name):
    if name == "er198":
        return _er_graph(gu)
    return gu.HostGraph.parse_metis(graph_path(name))


def _oracle_rows(orc, hg, sources):
    d = np.empty((len(sources), hg.n), dtype=np.int32)
    s = np.empty((len(sources), hg.n), dtype=np.uint64)
    for i, src in enumerate(sources):
        d[i], s[i] = orc.sigma(hg.R, hg.C, int(src))
    return d, s


def _search_and_compare(gu, orc, hg, sources, variant, direction, width, label, ref=None):
    announce("%s variant=%s dir=%s width=%d nsrc=%d n=%d" % (label, variant, direction, width,
                                                             len(sources), hg.n))
    cfg = gu.make_config(variant, DIRS[direction], width)
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(sources, dist=True, sigma=True, config=cfg)
    od, os_ = ref if ref is not None else _oracle_rows(orc, hg, sources)
    bad = np.argwhere(d != od)
    assert bad.size == 0, "distance mismatch at (row, v) = %s" % (bad[:5].tolist(),)
    bad = np.argwhere(s != os_)
    assert bad.size == 0, "path count mismatch at (row, v) = %s" % (bad[:5].tolist(),)
    return st


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("direction", ["auto", "topdown", "bottomup"])
def test_fixture_all_sources(gu, orc, gpu, name, variant, direction):
    hg = _load(gu, name)
    sources = np.arange(hg.n, dtype=np.int32)
    _search_and_compare(gu, orc, hg, sources, variant, direction, 64, name)


@pytest.mark.parametrize("width", [128, 256])
@pytest.mark.parametrize("variant", ["shuffle", "edge_parallel", "scan"])
def test_er_wide_batches(gu, orc, gpu, width, variant):
    hg = _er_graph(gu)
    sources = np.arange(hg.n, dtype=np.int32)
    _search_and_compare(gu, orc, hg, sources, variant, "auto", width, "er198")


@pytest.fixture(scope="module")
def rmat14(gu, orc):
    hg = gu.HostGraph.rmat(14, 16, seed=1)
    src = hg.sample_sources(64, seed=2)
    return hg, src, _oracle_rows(orc, hg, src)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("direction", ["auto", "topdown"])
def test_rmat14_64_sources(gu, orc, gpu, rmat14, variant, direction):
    hg, src, ref = rmat14
    _search_and_compare(gu, orc, hg, src, variant, direction, 64, "rmat14", ref=ref)


def test_rmat14_has_hubs(gu, rmat14):
    hg, _, _ = rmat14
    assert hg.degrees().max() > 1024


@pytest.mark.parametrize("width", [128, 256])
def test_rmat14_wide(gu, orc, gpu, rmat14, width):
    hg, src, _ = rmat14
    src2 = np.concatenate([src, hg.sample_sources(width - 64, seed=11)]).astype(np.int32)
    _search_and_compare(gu, orc, hg, src2, "edge_parallel", "auto", width, "rmat14")


def test_rmat14_duplicate_and_partial_batch(gu, orc, gpu, rmat14):
    hg, src, _ = rmat14
    sources = np.concatenate([src[:10], src[:10], src[3:40]]).astype(np.int32)  # 57 lanes, repeats
    _search_and_compare(gu, orc, hg, sources, "shuffle", "auto", 64, "rmat14-dups")


def test_grid_256_sources_and_diameter(gu, orc, gpu):
    hg = gu.HostGraph.grid(48, 0.10, seed=3)
    src = hg.sample_sources(256, seed=4)
    _search_and_compare(gu, orc, hg, src, "shuffle", "auto", 256, "grid48")
    _search_and_compare(gu, orc, hg, src[:64], "warp", "topdown", 64, "grid48")
    announce("grid48 diameter over 64 sources")
    with gu.DeviceGraph(hg) as dg:
        diam, st = dg.diameter(src[:64])
    assert diam == max(orc.eccentricity(hg.R, hg.C, int(s)) for s in src[:64])


@pytest.mark.parametrize("width", [64, 256])
def test_intact_grid_binomials_wrap(gu, gpu, width):
    """40x40 grid from a corner: sigma = C(r+c, r) mod 2^64 (wraps past
    2^64, which the overflow flag reports)."""
    from math import comb
    W = 40
    hg = gu.HostGraph.grid(W, 0.0, seed=0, lcc_only=False)
    src = np.array([0, W - 1, W * W - 1, 5 * W + 7] * (width // 4), dtype=np.int32)
    announce("intact grid 40x40 width=%d" % width)
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(src, True, True, gu.make_config("auto", 0, width))
    r, c = np.divmod(np.arange(W * W), W)
    exp = np.array([comb(int(a + b), int(a)) % (1 << 64) for a, b in zip(r, c)], dtype=np.uint64)
    assert np.array_equal(d[0], r + c)
    assert np.array_equal(s[0], exp)
    assert st.sigma_overflow == 1
    # corner W-1: distance r + (W-1-c)
    assert np.array_equal(d[1], r + (W - 1 - c))


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
def test_diameter_fixtures(gu, orc, gpu, name):
    hg = _load(gu, name)
    announce("diameter %s" % name)
    with gu.DeviceGraph(hg) as dg:
        diam, _ = dg.diameter(np.arange(hg.n, dtype=np.int32))
    assert diam == max(orc.eccentricity(hg.R, hg.C, s) for s in range(hg.n))


def _check_bc(gu, orc, hg, sources, width, label, variant="auto", engine=0):
    announce("bc %s width=%d nsrc=%d engine=%d" % (label, width, len(sources), engine))
    cfg = gu.make_config(variant, 0, width, engine=engine)
    with gu.DeviceGraph(hg) as dg:
        delta, bc, st = dg.betweenness(sources, delta=True, bc=True, config=cfg)
    if engine:
        assert st.engine_used == engine
    ref_bc = np.zeros(hg.n)
    for i, s in enumerate(sources):
        ref = orc.brandes_f64(hg.R, hg.C, int(s))
        ref_bc += ref
        nz = ref != 0
        assert np.all(delta[i][~nz] == 0), "delta should be 0 where FP64 delta is 0 (src %d)" % s
        rel = np.abs(delta[i][nz] - ref[nz]) / np.abs(ref[nz])
        assert rel.max(initial=0) <= BC_RTOL, "delta rel err %.3g (src %d)" % (rel.max(), s)
    nz = ref_bc != 0
    assert np.all(bc[~nz] == 0)
    rel = np.abs(bc[nz] - ref_bc[nz]) / np.abs(ref_bc[nz])
    assert rel.max(initial=0) <= BC_RTOL, "BC rel err %.3g" % rel.max()
    return st


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
def test_bc_fixtures(gu, orc, gpu, name):
    hg = _load(gu, name)
    _check_bc(gu, orc, hg, np.arange(hg.n, dtype=np.int32), 64, name)


def test_bc_closed_forms(gu, gpu):
    # unhalved BC over all sources: star center (n-1)(n-2), path 2i(n-1-i)
    with gu.DeviceGraph(_load(gu, "test2")) as dg:
        _, bc, _ = dg.betweenness(np.arange(6, dtype=np.int32), delta=False)
    assert bc[0] == 20 and np.all(bc[1:] == 0)
    with gu.DeviceGraph(_load(gu, "test5")) as dg:
        _, bc, _ = dg.betweenness(np.arange(7, dtype=np.int32), delta=False)
    assert np.allclose(bc, [0, 10, 16, 18, 16, 10, 0], rtol=0, atol=1e-9)


@pytest.mark.parametrize("width", [64, 128, 256])
def test_bc_rmat12(gu, orc, gpu, width):
    hg = gu.HostGraph.rmat(12, 16, seed=5)
    src = hg.sample_sources(200, seed=6)
    _check_bc(gu, orc, hg, src, width, "rmat12")


def _hub_graph(gu):
    """A 4000-vertex path plus two hubs (degrees 2000 and 1334) wired to
    every 2nd / 3rd path vertex: hub adjacency lists span several segments
    and the hubs sit on many shortest paths."""
    n = 4002
    edges = [(i, i + 1) for i in range(3999)]
    edges += [(4000, i) for i in range(0, 4000, 2)]
    edges += [(4001, i) for i in range(1, 4000, 3)]
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(a) for a in adj])
    C = np.array([c for a in adj for c in sorted(a)], dtype=np.int32)
    return gu.HostGraph(R, C)


@pytest.mark.parametrize("direction", ["auto", "topdown", "bottomup"])
@pytest.mark.parametrize("variant", ["shuffle", "edge_parallel"])
def test_hub_graph_search(gu, orc, gpu, direction, variant):
    hg = _hub_graph(gu)
    src = np.array([0, 1, 2, 3999, 2000, 4000, 4001, 1234] * 8, dtype=np.int32)
    _search_and_compare(gu, orc, hg, src, variant, direction, 64, "hubgraph")


def _hub_multigraph(gu):
    """The hub graph with parallel edges: source-to-hub edges doubled and
    tripled, so path counts into a hub at level 1 are edge multiplicities
    (the reference's adjacency walk counts every entry)."""
    hg = _hub_graph(gu)
    n = hg.n
    adj = [list(hg.C[hg.R[v]:hg.R[v + 1]]) for v in range(n)]
    for u, v, times in [(4000, 0, 1), (4001, 1, 2), (4000, 2, 1), (4001, 4000, 0), (4000, 4001, 1),
                        (3, 4000, 0)]:
        for _ in range(times):
            adj[u].append(v)
            adj[v].append(u)
    if 4001 not in adj[4000]:
        adj[4000].append(4001)
        adj[4001].append(4000)
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(a) for a in adj])
    C = np.array([c for a in adj for c in sorted(a)], dtype=np.int32)
    return gu.HostGraph(R, C)


@pytest.mark.parametrize("direction", ["auto", "topdown", "bottomup"])
@pytest.mark.parametrize("width", [64, 128])
def test_hub_multigraph_counts(gu, orc, gpu, direction, width):
    hg = _hub_multigraph(gu)
    src = np.array([0, 1, 2, 3, 4000, 4001, 3999, 7] * (width // 8), dtype=np.int32)
    _search_and_compare(gu, orc, hg, src, "shuffle", direction, width, "hubmultigraph")
    _, s1 = orc.sigma(hg.R, hg.C, 1)
    assert s1[4001] == 3  # three parallel edges from source 1


@pytest.mark.parametrize("width", [64, 256])
def test_bc_hub_graph(gu, orc, gpu, width):
    hg = _hub_graph(gu)
    src = np.arange(0, 4002, 37, dtype=np.int32)
    _check_bc(gu, orc, hg, src, width, "hubgraph")


def test_bc_rmat14_hubs(gu, orc, gpu, rmat14):
    hg, src, _ = rmat14
    assert hg.degrees().max() > 1024  # exercises the hub segments
    _check_bc(gu, orc, hg, src[:48], 64, "rmat14")


def test_bc_grid_deep(gu, orc, gpu):
    hg = gu.HostGraph.grid(24, 0.10, seed=7)
    src = hg.sample_sources(64, seed=8)
    _check_bc(gu, orc, hg, src, 64, "grid24")


def _directed_graph(gu, n=3000, arcs=20000, seed=12):
    """Random directed graph plus an in-hub (vertex 0, in-degree ~n/2) and an
    out-hub (vertex 1): pulls must use in-edges, expansion out-edges."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, arcs)
    dst = rng.integers(0, n, arcs)
    extra_src = np.concatenate([np.arange(2, n, 2), np.full(n // 2, 1)])
    extra_dst = np.concatenate([np.zeros(len(range(2, n, 2)), dtype=np.int64),
                                rng.integers(2, n, n // 2)])
    src = np.concatenate([src, extra_src])
    dst = np.concatenate([dst, extra_dst])
    keep = src != dst
    pairs = np.unique(np.stack([src[keep], dst[keep]], 1), axis=0)
    R = np.zeros(n + 1, dtype=np.int32)
    np.add.at(R, pairs[:, 0] + 1, 1)
    R = np.cumsum(R).astype(np.int32)
    C = pairs[:, 1].astype(np.int32)  # rows sorted by (src, dst)
    return gu.HostGraph(R, C, directed=True)


@pytest.mark.parametrize("direction", ["auto", "topdown", "bottomup"])
@pytest.mark.parametrize("variant", ["edge_parallel", "warp", "shuffle"])
@pytest.mark.parametrize("width", [64, 128])
def test_directed_graph(gu, orc, gpu, direction, variant, width):
    hg = _directed_graph(gu)
    src = np.concatenate([[0, 1, 2], hg.sample_sources(width - 3, seed=13)]).astype(np.int32)
    _search_and_compare(gu, orc, hg, src, variant, direction, width, "directed")


LBS_CASES = {
    "small": [0, 3, 0, 5, 1, 0, 2],
    "zeros_ends": [0, 0, 4, 0, 0, 0, 7, 1, 0, 0],
    "single": [17],
    "all_zero": [0, 0, 0],
    "one_item_each": [1] * 5000,
}


@pytest.mark.parametrize("case", sorted(LBS_CASES))
def test_lbs_cases(gu, orc, gpu, case):
    announce("lbs %s" % case)
    c = np.array(LBS_CASES[case], dtype=np.int32)
    scan, res, W = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW and np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))
    assert np.array_equal(res, np.searchsorted(oscan, np.arange(W), side="right") - 1)


@pytest.mark.parametrize("seed", [0, 1])
def test_lbs_random_with_zero_segments(gu, orc, gpu, seed):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, 300, size=20000).astype(np.int32)
    c[rng.random(c.size) < 0.1] = 0
    c[:3] = 0
    c[-3:] = 0
    announce("lbs random seed=%d" % seed)
    scan, res, W = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW and np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))


@pytest.mark.parametrize("S", [8191, 8192, 8193])
def test_lbs_single_launch_boundary(gu, orc, gpu, S):
    """Up to 8192 segments one launch scans and expands; one more takes the
    three-phase scan.  Both sides of the boundary give the oracle's output."""
    rng = np.random.default_rng(S)
    c = rng.integers(0, 40, size=S).astype(np.int32)
    c[rng.random(S) < 0.2] = 0
    announce("lbs boundary S=%d" % S)
    scan, res, W = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW and np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))


class _PtrView:
    """A device address inside another buffer (for the ctypes entry points)."""

    def __init__(self, ptr):
        self.ptr = ptr


@pytest.mark.parametrize("S", [3000, 20000])
@pytest.mark.parametrize("off", [1, 2, 3])
def test_lbs_unaligned_result(gu, orc, gpu, off, S):
    """A result pointer that is not 16-byte aligned takes one-item stores
    (aligned ones store four items at a time), in the one-launch form and
    beyond 8192 segments; the items either side of the result stay
    untouched."""
    import ctypes

    rng = np.random.default_rng(40 + off + S)
    c = rng.integers(0, 50, size=S).astype(np.int32)
    c[rng.random(c.size) < 0.1] = 0
    W = int(c.sum())
    announce("lbs unaligned result S=%d off=%d W=%d" % (S, off, W))
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * c.size)
    rb = gu.DeviceBuffer.from_array(np.full(W + 8, -7, dtype=np.int32))
    try:
        view = _PtrView(ctypes.c_void_p(rb.ptr.value + 4 * off))
        Wd, _ = gu.lbs_device(cb, c.size, sb, view, W)
        full = rb.to_host(np.int32)
    finally:
        for b in (cb, sb, rb):
            b.close()
    oscan, oW = orc.exclusive_scan(c)
    assert Wd == oW
    assert np.array_equal(full[off:off + W], orc.lbs(oW, oscan))
    assert (full[:off] == -7).all() and (full[off + W:] == -7).all()


def test_lbs_total_overflow_is_capacity_error(gu, gpu):
    """A total of 2^31 + 2^30 items exceeds the int32 output index: refused
    with a capacity error and nothing written."""
    c = np.array([1 << 30, 1 << 30, 1 << 30], dtype=np.int32)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * 3)
    rb = gu.DeviceBuffer(4 * 16)
    rb.zero()
    announce("lbs total overflow")
    with pytest.raises(gu.GuError) as ei:
        gu.lbs_device(cb, 3, sb, rb, 16)
    assert ei.value.status == gu.ERR_CAPACITY
    assert np.all(rb.to_host(np.int32) == 0)


@pytest.mark.parametrize("S", [100000, 1000000])
def test_lbs_many_zipf_segments(gu, orc, gpu, S):
    """The reference's own LBS example size (100,000 segments) and ten times
    that: tiles that cross many small segments stage them in LDS, or search
    the scan when more than the window; most tiles lie in one segment."""
    rng = np.random.default_rng(S)
    w = 1.0 / np.arange(1, S + 1) ** 1.1
    c = np.floor(w / w.sum() * (1 << 22)).astype(np.int64)
    c[0] += (1 << 22) - c.sum()
    rng.shuffle(c)
    c = c.astype(np.int32)
    announce("lbs zipf S=%d W=%d" % (S, int(c.sum())))
    scan, res, W = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW and np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))


def test_lbs_zipf_2_22(gu, orc, gpu):
    c = gu.zipf_counts(1.1, 1 << 20, 1 << 22, seed=9)
    announce("lbs zipf W=2^22 S=%d" % len(c))
    scan, res, W = gu.lbs(c)
    assert W == 1 << 22
    oscan, oW = orc.exclusive_scan(c)
    assert np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))


def test_lbs_capacity_error(gu, gpu):
    c = np.array([5, 5], dtype=np.int32)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(8)
    rb = gu.DeviceBuffer(4 * 4)
    rb.zero()
    announce("lbs capacity error")
    with pytest.raises(gu.GuError) as ei:
        gu.lbs_device(cb, 2, sb, rb, 4)
    assert ei.value.status == gu.ERR_CAPACITY
    assert np.all(rb.to_host(np.int32) == 0)  # nothing written


@pytest.mark.parametrize("S", [5000, 300000])
def test_lbs_scan_only(gu, orc, gpu, S):
    """gu_lbs with no result buffer computes the scan and the total only and
    returns GU_OK, on the one-launch path (up to 8192 segments) and the
    three-phase scan."""
    c = np.random.default_rng(S).integers(0, 50, S).astype(np.int32)
    cb = gu.DeviceBuffer.from_array(c)
    sb = gu.DeviceBuffer(4 * S)
    announce("lbs scan only S=%d" % S)
    try:
        W, _ = gu.lbs_device(cb, S, sb, None, 0)
        scan = sb.to_host(np.int32, S)
    finally:
        cb.close()
        sb.close()
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW and np.array_equal(scan, oscan)


def test_source_out_of_range(gu, gpu):
    hg = _load(gu, "test1")
    announce("source out of range")
    with gu.DeviceGraph(hg) as dg:
        with pytest.raises(gu.GuError) as ei:
            dg.multi_search(np.array([0, 9], dtype=np.int32))
        assert ei.value.status == gu.ERR_RANGE


def test_repeatable(gu, gpu, rmat14):
    hg, src, _ = rmat14
    announce("repeatability rmat14")
    with gu.DeviceGraph(hg) as dg:
        a = dg.multi_search(src, True, True)
        b = dg.multi_search(src, True, True)
        c1 = dg.betweenness(src[:32])
        c2 = dg.betweenness(src[:32])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(c1[0], c2[0]) and np.array_equal(c1[1], c2[1])


# ---- per-source engine (one workgroup per source, level loop in one launch)

PS = 2  # GU_ENGINE_PER_SOURCE


def _ps_compare(gu, orc, hg, sources, label, sigma=True):
    announce("per-source %s nsrc=%d n=%d" % (label, len(sources), hg.n))
    cfg = gu.make_config(engine=PS)
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(sources, dist=True, sigma=sigma, config=cfg)
    assert st.engine_used == PS
    od, os_ = _oracle_rows(orc, hg, sources)
    bad = np.argwhere(d != od)
    assert bad.size == 0, "distance mismatch at %s" % (bad[:5].tolist(),)
    if sigma:
        bad = np.argwhere(s != os_)
        assert bad.size == 0, "path count mismatch at %s" % (bad[:5].tolist(),)
    return st


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
def test_persource_fixtures(gu, orc, gpu, name):
    hg = _load(gu, name)
    _ps_compare(gu, orc, hg, np.arange(hg.n, dtype=np.int32), name)


def test_persource_rmat14_and_hubs(gu, orc, gpu, rmat14):
    hg, src, _ = rmat14
    _ps_compare(gu, orc, hg, src[:32], "rmat14")
    _ps_compare(gu, orc, _hub_graph(gu), np.array([0, 4000, 4001, 3999], dtype=np.int32), "hubgraph")


def test_persource_grid_and_diameter(gu, orc, gpu):
    hg = gu.HostGraph.grid(48, 0.10, seed=3)
    src = hg.sample_sources(300, seed=4)  # more than one 256-source launch: the queue-fed grid
    st = _ps_compare(gu, orc, hg, src, "grid48")
    assert st.batches == 1
    with gu.DeviceGraph(hg) as dg:
        diam, st = dg.diameter(src[:64], gu.make_config(engine=PS))
    assert st.engine_used == PS
    assert diam == max(orc.eccentricity(hg.R, hg.C, int(s)) for s in src[:64])


def test_persource_directed(gu, orc, gpu):
    hg = _directed_graph(gu)
    _ps_compare(gu, orc, hg, np.concatenate([[0, 1, 2], hg.sample_sources(29, seed=5)]).astype(np.int32),
                "directed")


def _directed_degree4(gu, n=5000, seed=77):
    """Random directed graph of out-degree 0-4, no self-loops or parallel
    arcs."""
    rng = np.random.default_rng(seed)
    rows = []
    for v in range(n):
        k = int(rng.integers(0, 5))
        t = rng.choice(n - 1, size=k, replace=False)
        rows.append(np.sort(np.where(t >= v, t + 1, t)))
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(r) for r in rows])
    C = np.concatenate(rows).astype(np.int32)
    hg = gu.HostGraph(R, C, directed=True)
    assert hg.degrees().max() <= 4
    return hg


def test_persource_directed_degree4(gu, orc, gpu):
    """Directed graph of out-degree <= 4 (some vertices without out-edges):
    the per-source kernel reads it as one int4 of neighbours per vertex and
    pushes path counts along arcs."""
    hg = _directed_degree4(gu)
    _ps_compare(gu, orc, hg, hg.sample_sources(64, seed=78), "directed-deg4")
    _ps_compare(gu, orc, hg, hg.sample_sources(16, seed=79), "directed-deg4-d", sigma=False)


def _kary_tree(gu, k, depth):
    """Complete k-ary tree of the given depth (heap numbering): its deepest
    level holds k^(depth-1) vertices, more than the per-source LDS frontier,
    so the HBM spill path runs."""
    n = (k ** depth - 1) // (k - 1)
    adj = [[] for _ in range(n)]
    for v in range(1, n):
        p = (v - 1) // k
        adj[p].append(v)
        adj[v].append(p)
    R = np.zeros(n + 1, dtype=np.int32)
    R[1:] = np.cumsum([len(a) for a in adj])
    C = np.array([c for a in adj for c in sorted(a)], dtype=np.int32)
    return gu.HostGraph(R, C)


@pytest.mark.parametrize("k,depth", [(2, 15), (4, 7)])
def test_persource_tree_frontier_spill(gu, orc, gpu, k, depth):
    """Binary tree (degree-4 neighbour table): 16384 leaves; 4-ary tree
    (degree 5, adjacency bounds in the frontier): 4096.  Both exceed the
    3072-entry LDS frontier."""
    hg = _kary_tree(gu, k, depth)
    src = np.array([0, 1, 2, hg.n - 1, hg.n // 2, 77], dtype=np.int32)
    _ps_compare(gu, orc, hg, src, "tree%d-%d" % (k, depth))
    _check_bc(gu, orc, hg, src, 64, "tree%d-%d" % (k, depth), engine=PS)


def _undirected_csr(gu, n, edges):
    """Undirected HostGraph from an (E, 2) edge array: both directions, rows
    sorted, duplicates and self-loops dropped."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    both = np.unique(np.concatenate([e, e[:, ::-1]]), axis=0)
    R = np.zeros(n + 1, dtype=np.int32)
    np.add.at(R, both[:, 0] + 1, 1)
    R = np.cumsum(R).astype(np.int32)
    return gu.HostGraph(R, both[:, 1].astype(np.int32))


def _facts(hg):
    """(bipartite, connected) by a 2-colouring BFS over every component."""
    R, C, n = hg.R, hg.C, hg.n
    col = np.full(n, -1, dtype=np.int8)
    bip, comps = True, 0
    for s in range(n):
        if col[s] >= 0:
            continue
        comps += 1
        col[s] = 0
        q = [s]
        for v in q:
            for w in C[R[v]:R[v + 1]]:
                if col[w] < 0:
                    col[w] = col[v] ^ 1
                    q.append(int(w))
                elif col[w] == col[v]:
                    bip = False
    return bip, comps == 1


def _lds_kind_graph(gu, kind):
    rng = np.random.default_rng(91)
    if kind == "odd_cycle":
        n = 2001
        return _undirected_csr(gu, n, np.stack([np.arange(n), (np.arange(n) + 1) % n], 1))
    if kind == "grid_components":
        return gu.HostGraph.grid(64, 0.30, seed=92, lcc_only=False)
    if kind in ("random_deg4", "random_deg4_wide"):
        n = 20000 if kind == "random_deg4" else 120000  # the wide one: neighbour deltas beyond int16
        stubs = rng.permutation(np.repeat(np.arange(n), rng.integers(0, 5, n)))  # degrees 0-4
        return _undirected_csr(gu, n, stubs[: len(stubs) // 2 * 2].reshape(-1, 2))
    if kind == "grid_diagonals":
        W = 60
        r, c = np.divmod(np.arange(W * W), W)
        v = np.arange(W * W)
        e = [np.stack([v[c < W - 1], v[c < W - 1] + 1], 1), np.stack([v[r < W - 1], v[r < W - 1] + W], 1)]
        # a diagonal across every 6th interior cell joins two vertices of one
        # colour (odd cycles); the cell's top and bottom edges go, so no
        # vertex exceeds degree 4
        keep = set(map(tuple, np.concatenate(e).tolist()))
        for a in range(1, W - 2, 3):
            for b in range(1, W - 2, 2):
                x = a * W + b
                keep.discard((x, x + 1))
                keep.discard((x + W, x + W + 1))
                keep.add((x, x + W + 1))
        return _undirected_csr(gu, W * W, np.array(sorted(keep)))
    raise ValueError(kind)


@pytest.mark.parametrize("kind,facts", [("odd_cycle", (False, True)), ("grid_components", (True, False)),
                                        ("random_deg4", (False, False)), ("random_deg4_wide", (False, False)),
                                        ("grid_diagonals", (False, True))])
def test_persource_lds_graph_kinds(gu, orc, gpu, kind, facts):
    """Undirected degree <= 4 graphs of each kind the level-set search
    treats apart (bipartite or not: a same-level lookup or none; connected
    or not: a distance-row fill or none; neighbour ids within int16 of the
    vertex's own or not: the compact neighbour table or the int4 one),
    against the oracle: rows, path counts and dependencies."""
    hg = _lds_kind_graph(gu, kind)
    assert hg.degrees().max() <= 4
    assert _facts(hg) == facts
    for k, seed in ((45, 93), (7, 94)):
        _ps_compare(gu, orc, hg, hg.sample_sources(k, seed=seed), kind)
    _ps_compare(gu, orc, hg, hg.sample_sources(9, seed=95), kind + "-d", sigma=False)
    _check_bc(gu, orc, hg, hg.sample_sources(45, seed=96), 64, kind, engine=PS)


@pytest.mark.parametrize("sigma", [True, False])
def test_persource_path_deeper_than_16bit_levels(gu, orc, gpu, sigma):
    """A path of 70000 vertices searched from an end: 69999 levels, past
    what the level-set search's 16-bit distance rows hold (65534), so the
    launch is rerun on the distance kernel; rows and counts still match,
    with path counts and for distances only."""
    n = 70000
    hg = _undirected_csr(gu, n, np.stack([np.arange(n - 1), np.arange(1, n)], 1))
    st = _ps_compare(gu, orc, hg, np.array([0, n - 1, n // 2, 3], dtype=np.int32), "path70000", sigma=sigma)
    assert st.max_level == n - 1


@pytest.mark.parametrize("side", [2896, 2897])
def test_persource_level_set_key_range(gu, orc, gpu, side):
    """Intact grids on either side of the id range of the keyed level sets
    (2^23 = 8,388,608 vertices), which a search with path counts uses:
    2896^2 = 8,386,816 vertices run them with ids near the top of the range;
    2897^2 = 8,392,609 take the distance kernel for path counts, and plain
    level sets (no range limit) for distances only.  Corner, centre and
    interior sources against the oracle (path counts wrap modulo 2^64 on an
    intact grid)."""
    hg = gu.HostGraph.grid(side, 0.0, seed=1)
    assert hg.n == side * side and hg.degrees().max() <= 4
    n = hg.n
    src = np.array([0, n - 1, n // 2, 12345], dtype=np.int32)
    _ps_compare(gu, orc, hg, src, "grid%d" % side)
    _ps_compare(gu, orc, hg, src[:2], "grid%d-d" % side, sigma=False)


def test_persource_diameter_on_level_sets(gu, orc, gpu):
    """Diameter sampling (eccentricities only) runs the level-set search
    without rows; sources whose search overflows (the 70000-vertex path's
    ends, deeper than 16-bit levels; a binary tree's 16384-vertex level) are
    searched again on the distance kernel."""
    n = 70000
    path = _undirected_csr(gu, n, np.stack([np.arange(n - 1), np.arange(1, n)], 1))
    tree = _kary_tree(gu, 2, 15)
    grid = gu.HostGraph.grid(48, 0.10, seed=3)
    cases = [(path, np.array([n // 2, 0, 17], dtype=np.int32), "path70000"),
             (tree, np.array([0, 5, tree.n - 1], dtype=np.int32), "tree2-15"),
             (grid, grid.sample_sources(40, seed=8), "grid48")]
    for hg, src, label in cases:
        announce("diameter on level sets %s nsrc=%d" % (label, len(src)))
        with gu.DeviceGraph(hg) as dg:
            diam, st = dg.diameter(src, gu.make_config(engine=PS))
            ecc, _ = dg.multi_search_rows(src, None, None, want_ecc=True, config=gu.make_config(engine=PS))
        assert st.engine_used == PS
        expect = [orc.eccentricity(hg.R, hg.C, int(s)) for s in src]
        assert list(ecc) == expect, label
        assert diam == max(expect), label


def test_persource_distances_only(gu, orc, gpu):
    hg = gu.HostGraph.grid(32, 0.10, seed=9)
    _ps_compare(gu, orc, hg, hg.sample_sources(40, seed=10), "grid32-d", sigma=False)


def test_persource_intact_grid_wrap(gu, gpu):
    from math import comb
    W = 40
    hg = gu.HostGraph.grid(W, 0.0, seed=0, lcc_only=False)
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(np.array([0], dtype=np.int32), True, True, gu.make_config(engine=PS))
    r, c = np.divmod(np.arange(W * W), W)
    assert np.array_equal(d[0], r + c)
    assert np.array_equal(s[0], np.array([comb(int(a + b), int(a)) % (1 << 64) for a, b in zip(r, c)],
                                         dtype=np.uint64))
    assert st.sigma_overflow == 1


@pytest.mark.parametrize("name", FIXTURES + ["er198"])
def test_persource_bc_fixtures(gu, orc, gpu, name):
    hg = _load(gu, name)
    _check_bc(gu, orc, hg, np.arange(hg.n, dtype=np.int32), 64, name, engine=PS)


def test_persource_bc_grid_two_launches(gu, orc, gpu):
    hg = gu.HostGraph.grid(30, 0.10, seed=31)
    _check_bc(gu, orc, hg, hg.sample_sources(300, seed=32), 64, "grid30", engine=PS)


def test_persource_bc_hubs(gu, orc, gpu):
    _check_bc(gu, orc, _hub_graph(gu), np.arange(0, 4002, 97, dtype=np.int32), 64, "hubgraph", engine=PS)


def _brandes_directed_f64(R, C, n, s):
    """Brandes on a directed graph (dependencies pulled over out-edges).
    The reference's sequential Brandes (bc.cu:11-57) treats the adjacency as
    undirected, so directed graphs are checked against this textbook form."""
    from collections import deque
    d = np.full(n, -1)
    sg = np.zeros(n)
    d[s], sg[s] = 0, 1.0
    order, q = [], deque([s])
    while q:
        v = q.popleft()
        order.append(v)
        for w in C[R[v]:R[v + 1]]:
            if d[w] < 0:
                d[w] = d[v] + 1
                q.append(w)
            if d[w] == d[v] + 1:
                sg[w] += sg[v]
    delta = np.zeros(n)
    for v in reversed(order):
        for w in C[R[v]:R[v + 1]]:
            if d[w] == d[v] + 1:
                delta[v] += sg[v] / sg[w] * (1.0 + delta[w])
    delta[s] = 0.0
    return delta


@pytest.mark.parametrize("engine,graph", [(1, "hubs"), (PS, "hubs"), (PS, "degree4")])
def test_bc_directed(gu, gpu, engine, graph):
    hg = _directed_graph(gu) if graph == "hubs" else _directed_degree4(gu, 3000, 91)
    src = np.concatenate([[0, 1], hg.sample_sources(30, seed=7)]).astype(np.int32)
    announce("bc directed engine=%d" % engine)
    with gu.DeviceGraph(hg) as dg:
        delta, bc, st = dg.betweenness(src, config=gu.make_config(engine=engine))
    assert st.engine_used == engine
    ref_bc = np.zeros(hg.n)
    for i, s in enumerate(src):
        ref = _brandes_directed_f64(hg.R, hg.C, hg.n, int(s))
        ref_bc += ref
        nz = ref != 0
        assert np.all(delta[i][~nz] == 0)
        assert np.max(np.abs(delta[i][nz] - ref[nz]) / ref[nz], initial=0) <= BC_RTOL
    nz = ref_bc != 0
    assert np.all(bc[~nz] == 0)
    assert np.max(np.abs(bc[nz] - ref_bc[nz]) / ref_bc[nz]) <= BC_RTOL


def test_persource_bc_matches_batched(gu, gpu):
    """Same FP32 recurrence in the same adjacency order: on a hub-free graph
    the two engines' dependency rows agree bit for bit."""
    hg = gu.HostGraph.grid(40, 0.10, seed=33)
    src = hg.sample_sources(100, seed=34)
    with gu.DeviceGraph(hg) as dg:
        d1, b1, _ = dg.betweenness(src, config=gu.make_config(engine=1))
        d2, b2, _ = dg.betweenness(src, config=gu.make_config(engine=PS))
    assert np.array_equal(d1, d2)
    assert np.array_equal(b1, b2)


def test_engine_auto_choice(gu, gpu, rmat14):
    hg, src, _ = rmat14
    with gu.DeviceGraph(hg) as dg:
        _, _, st = dg.multi_search(src[:8], True, False)
    assert st.engine_used == 1  # batched on a power-law graph
    hg = gu.HostGraph.grid(32, 0.10, seed=9)
    with gu.DeviceGraph(hg) as dg:
        _, _, st = dg.multi_search(hg.sample_sources(8, seed=1), True, False)
    assert st.engine_used == PS


@pytest.mark.parametrize("graphs", [1, 0])
def test_level_graph_replay_matches_direct(gu, orc, gpu, graphs):
    """Captured chunks of levels replayed many times (several batches, widths
    and both job kinds, with eager work in between) equal direct launches and
    the oracle."""
    hg = gu.HostGraph.rmat(12, 16, seed=51)
    announce("level graphs=%d" % graphs)
    with gu.DeviceGraph(hg) as dg:
        for rep in range(3):
            for width in (64, 128):
                src = hg.sample_sources(width, seed=100 * rep + width)
                d, s, _ = dg.multi_search(src, True, True, gu.make_config(level_graphs=graphs))
                od, os_ = _oracle_rows(orc, hg, src[:6])
                assert np.array_equal(d[:6], od) and np.array_equal(s[:6], os_)
                d2, s2, _ = dg.multi_search(src, True, True, gu.make_config(level_graphs=0))
                assert np.array_equal(d, d2) and np.array_equal(s, s2)
            x1, b1, _ = dg.betweenness(src[:40], config=gu.make_config(level_graphs=graphs))
            x2, b2, _ = dg.betweenness(src[:40], config=gu.make_config(level_graphs=0))
            assert np.array_equal(x1, x2) and np.array_equal(b1, b2)


# ---- randomized sweep: random graphs x random launch shapes ---------------

@pytest.mark.parametrize("seed", range(12))
def test_random_graphs_and_shapes(gu, orc, gpu, seed):
    """Seeded random graphs (ER, R-MAT, grid, directed) and random source
    sets, engines, expanders, widths, workgroup sizes and direction rules:
    distances and path counts bit-exact, BC within 1e-5 of FP64."""
    rng = np.random.default_rng(1000 + seed)
    kind = ["er", "rmat", "grid", "directed"][seed % 4]
    if kind == "er":
        hg = gu.HostGraph.er(int(rng.integers(50, 600)), int(rng.integers(100, 3000)), seed=seed)
    elif kind == "rmat":
        hg = gu.HostGraph.rmat(int(rng.integers(8, 12)), int(rng.integers(4, 17)), seed=seed)
    elif kind == "grid":
        hg = gu.HostGraph.grid(int(rng.integers(8, 40)), float(rng.uniform(0, 0.3)), seed=seed)
    else:
        hg = _directed_graph(gu, n=int(rng.integers(200, 1500)), arcs=int(rng.integers(1000, 8000)),
                             seed=seed)
    k = int(rng.integers(1, min(300, hg.n) + 1))
    src = rng.integers(0, hg.n, k).astype(np.int32)
    variant = ["linear", "edge_parallel", "warp", "scan", "cta_warp", "shuffle", "auto"][int(rng.integers(7))]
    cfg = gu.make_config(variant, direction=int(rng.integers(3)),
                         batch_width=[0, 64, 128, 256][int(rng.integers(4))],
                         wg_size=[0, 256, 512, 1024][int(rng.integers(4))],
                         alpha=[0, 2, 16, 256][int(rng.integers(4))],
                         beta=[0, 8, 64][int(rng.integers(3))],
                         engine=int(rng.integers(3)))
    announce("random %s seed=%d n=%d k=%d variant=%s engine=%d" % (kind, seed, hg.n, k, variant, cfg.engine))
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(src, True, True, cfg)
        sub = src[:16]
        od, os_ = _oracle_rows(orc, hg, sub)
        assert np.array_equal(d[:len(sub)], od)
        assert np.array_equal(s[:len(sub)], os_)
        if not hg.directed:
            bsub = src[:24]
            _, bc, _ = dg.betweenness(bsub, delta=False, config=cfg)
            ref = orc.bc_f64(hg.R, hg.C, bsub)
            nz = ref > 0
            assert np.all(bc[~nz] == 0)
            if nz.any():
                assert np.max(np.abs(bc[nz] - ref[nz]) / ref[nz]) <= BC_RTOL
        diam, _ = dg.diameter(src[:32], cfg)
        assert diam == max(orc.eccentricity(hg.R, hg.C, int(x)) for x in src[:32])


@pytest.mark.parametrize("graph,engine,sigma", [("rmat12", 1, True), ("rmat12", 1, False), ("rmat12", PS, True),
                                                ("grid40", 1, True), ("grid40", PS, False),
                                                ("directed", 1, True), ("directed", PS, True)])
def test_teps_units_and_reached(gu, orc, gpu, graph, engine, sigma):
    """The GTEPS numerator: teps_units = sum over sources of the out-degrees
    of the vertices each reaches (shortest_paths.cu / multi_search_test.cu
    convention), and reached = sum over sources of reached vertices, as the
    oracle counts them from its distances."""
    if graph == "rmat12":
        hg = gu.HostGraph.rmat(12, 16, seed=12)
    elif graph == "grid40":
        hg = gu.HostGraph.grid(40, 0.1, seed=40)
    else:
        hg = _directed_graph(gu)
    src = hg.sample_sources(70, seed=3)
    announce("teps units %s engine=%d sigma=%d" % (graph, engine, sigma))
    with gu.DeviceGraph(hg) as dg:
        d, _, st = dg.multi_search(src, dist=True, sigma=sigma, config=gu.make_config(engine=engine))
        _, _, stb = dg.betweenness(src, delta=False, config=gu.make_config(engine=engine))
    units = sum(orc.teps_units(hg.R, orc.sigma(hg.R, hg.C, int(s))[0]) for s in src)
    reached = int(np.sum(d != np.iinfo(np.int32).max))
    assert st.engine_used == engine
    assert st.teps_units == units and st.reached == reached
    assert stb.teps_units == units and stb.reached == reached
