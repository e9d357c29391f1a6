"""GPU parity at the benchmark configurations' own sizes.

Each test runs the benchmark's own seeded input (same generator, seeds and
source samples as bench.py) through the C ABI and checks every output row:

- R-MAT scale 20, 64 sources, every top-down expander, auto and top-down
  only (multi_search_test.cu:40-127);
- the 2048x2048 grid (10% of edges deleted, largest component) with 256
  sources on the per-source engine, 64 on the batched engine, and 64-source
  diameter sampling (diameter_sampling_test.cu:40-72);
- sampled Brandes BC on R-MAT scale 22 over 1024 sources at 64/128/256
  sources per launch (bc.cu:59-146);
- the load-balanced search over 2^26 items (lbs.cu:87-97);
- the query server over the three resident graphs of the serving workload.

Rows are checked two ways.  A subset is compared bit for bit with the
oracle's rows.  Every row is checked with or_check_search_rows, whose BFS
equations (d[s] = 0, d[v] = 1 + min over in-neighbours, unreached vertices
have only unreached in-neighbours, sigma[v] = sum over the in-neighbours one
level closer) have exactly one solution, the oracle's rows; the checker
itself is pinned with planted mistakes in tests/test_oracle.py.

Brandes: BC over 64 of the sources within 1e-5 relative of the FP64 oracle
(exactly 0 where the oracle's is 0); for all 1024 sources the dependency rows
satisfy sum_v delta_s(v) = sum over reached t != s of (d_s(t) - 1) (the mean
number of interior vertices of a shortest s-t path, summed over t), with d_s
from searches that pass the checker; BC equals the sum of the rows, and is 0
on every vertex of degree <= 1.
"""
import numpy as np
import pytest
import xxhash

from conftest import announce

pytestmark = pytest.mark.gpu

VARIANTS = ["linear", "edge_parallel", "warp", "scan", "cta_warp", "shuffle"]
DIRS = {"auto": 0, "topdown": 1}
INT_MAX = 2147483647
BC_RTOL = 1e-5


def _check_rows(orc, hg, src, d, s, label):
    viol, where = orc.check_search_rows(hg.R, hg.C, src, d, s)
    assert viol == 0, "%s: %d violations of the BFS equations, first at (row, v) = %s" % (label, viol, where)


def _teps(hg, d):
    deg = np.diff(hg.R).astype(np.int64)
    return int(sum(deg[row != INT_MAX].sum() for row in d))


# ---------------------------------------------------------------------------
# R-MAT scale 20, 64 sources per batch (the default benchmark line)


@pytest.fixture(scope="module")
def rmat20(gu, orc, gpu):
    announce("generate R-MAT s20 (seed 20) and oracle rows")
    hg = gu.HostGraph.rmat(20, 16, 0.57, 0.19, 0.19, seed=20)
    src = hg.sample_sources(64, seed=1000)
    sub = np.arange(0, 64, 8)
    od, os_, _ = orc.rows_mt(hg.R, hg.C, src[sub], sigma=True)
    dg = gu.DeviceGraph(hg)
    yield hg, dg, src, sub, od, os_
    dg.close()


@pytest.mark.parametrize("direction", ["auto", "topdown"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_c2_rmat20_64_sources(gu, orc, gpu, rmat20, variant, direction):
    hg, dg, src, sub, od, os_ = rmat20
    announce("C2 rmat20 64 sources variant=%s dir=%s" % (variant, direction))
    cfg = gu.make_config(variant, DIRS[direction], 64)
    d, s, st = dg.multi_search(src, dist=True, sigma=True, config=cfg)
    assert st.variant_used == gu.VARIANTS[variant]
    if direction == "topdown":
        assert st.bottomup_levels == 0
    bad = np.argwhere(d[sub] != od)
    assert bad.size == 0, "distance mismatch at (row, v) = %s" % (bad[:5].tolist(),)
    bad = np.argwhere(s[sub] != os_)
    assert bad.size == 0, "path count mismatch at (row, v) = %s" % (bad[:5].tolist(),)
    _check_rows(orc, hg, src, d, s, "rmat20 %s %s" % (variant, direction))
    assert st.teps_units == _teps(hg, d)


# ---------------------------------------------------------------------------
# 2048^2 grid, 10% of the edges deleted, largest component


@pytest.fixture(scope="module")
def grid2048(gu, orc, gpu):
    announce("generate 2048^2 grid (seed 30) and oracle rows")
    hg = gu.HostGraph.grid(2048, 0.10, seed=30)
    src = hg.sample_sources(256, seed=3000)
    sub = np.arange(0, 256, 32)
    od, os_, _ = orc.rows_mt(hg.R, hg.C, src[sub], sigma=True)
    dg = gu.DeviceGraph(hg)
    yield hg, dg, src, sub, od, os_
    dg.close()


def _grid_check(gu, orc, hg, dg, src, sub, od, os_, engine, label):
    k, n = len(src), hg.n
    db, sb = gu.DeviceBuffer(4 * k * n), gu.DeviceBuffer(8 * k * n)
    try:
        st = dg.multi_search_device(src, db, sb, gu.make_config(engine=engine))
        assert st.engine_used == engine
        teps = 0
        for c0 in range(0, k, 64):
            c1 = min(k, c0 + 64)
            d = db.to_host(np.int32, (c1 - c0) * n, c0 * n).reshape(c1 - c0, n)
            s = sb.to_host(np.uint64, (c1 - c0) * n, c0 * n).reshape(c1 - c0, n)
            _check_rows(orc, hg, src[c0:c1], d, s, "%s rows %d-%d" % (label, c0, c1))
            for j, i in enumerate(sub):
                if c0 <= i < c1:
                    assert np.array_equal(d[i - c0], od[j]), "%s: distance row %d" % (label, i)
                    assert np.array_equal(s[i - c0], os_[j]), "%s: path count row %d" % (label, i)
            teps += _teps(hg, d)
        assert st.teps_units == teps
    finally:
        db.close()
        sb.close()


def test_c3_grid2048_256_sources_per_source_engine(gu, orc, gpu, grid2048):
    hg, dg, src, sub, od, os_ = grid2048
    announce("C3 grid2048 256 sources, per-source engine")
    _grid_check(gu, orc, hg, dg, src, sub, od, os_, gu.ENGINE_PER_SOURCE, "grid2048 per-source")


def test_c3_grid2048_64_sources_batched_engine(gu, orc, gpu, grid2048):
    hg, dg, src, sub, od, os_ = grid2048
    announce("C3 grid2048 64 sources, batched engine")
    _grid_check(gu, orc, hg, dg, src[:64], sub[sub < 64], od[: (sub < 64).sum()], os_[: (sub < 64).sum()],
                gu.ENGINE_BATCHED, "grid2048 batched")


@pytest.mark.parametrize("engine", [2, 1])
def test_c3_grid2048_diameter_64_sources(gu, orc, gpu, grid2048, engine):
    hg, dg, _, _, _, _ = grid2048
    dsrc = hg.sample_sources(64, seed=3999)
    announce("C3 grid2048 diameter, 64 sources, engine %d" % engine)
    cfg = gu.make_config(engine=engine)
    diam, st = dg.diameter(dsrc, config=cfg)
    ecc, _ = dg.multi_search_rows(dsrc, None, None, want_ecc=True, config=cfg)
    oecc = orc.eccentricities_mt(hg.R, hg.C, dsrc)
    assert np.array_equal(ecc, oecc)
    assert diam == int(oecc.max())


# ---------------------------------------------------------------------------
# Sampled Brandes BC on R-MAT scale 22, 1024 sources


@pytest.fixture(scope="module")
def rmat22(gu, orc, gpu):
    announce("generate R-MAT s22 (seed 22); FP64 oracle BC over 64 sources")
    hg = gu.HostGraph.rmat(22, 16, 0.57, 0.19, 0.19, seed=22)
    src = hg.sample_sources(1024, seed=2200)
    ref64 = orc.bc_f64_mt(hg.R, hg.C, src[:64], orc.default_threads())
    dg = gu.DeviceGraph(hg)
    # sum over reached t != s of (d_s(t) - 1), from distance rows that pass
    # the BFS-equation check
    dsum = np.zeros(len(src), dtype=np.float64)
    for c0 in range(0, len(src), 64):
        announce("s22 distance rows %d-%d" % (c0, c0 + 64))
        d, _, _ = dg.multi_search(src[c0:c0 + 64], dist=True, sigma=False, config=gu.make_config(batch_width=64))
        _check_rows(orc, hg, src[c0:c0 + 64], d, None, "rmat22 distances %d" % c0)
        fin = d != INT_MAX
        dsum[c0:c0 + 64] = (np.where(fin, d, 0).astype(np.int64).sum(axis=1) - fin.sum(axis=1) + 1)
    yield hg, dg, src, ref64, dsum
    dg.close()


@pytest.mark.parametrize("width", [64, 128, 256])
def test_c4a_bc_rmat22_1024_sources(gu, orc, gpu, rmat22, width):
    hg, dg, src, ref64, dsum = rmat22
    n, k = hg.n, len(src)
    cfg = gu.make_config(batch_width=width)
    announce("C4a rmat22 BC, 64 sources vs FP64, width %d" % width)
    _, bc64, st = dg.betweenness(src[:64], delta=False, bc=True, config=cfg)
    assert st.batch_width_used == width
    pos = ref64 > 0
    assert np.all(bc64[~pos] == 0), "BC nonzero where the FP64 oracle's is 0"
    err = np.abs(bc64[pos] - ref64[pos]) / ref64[pos]
    assert err.max() <= BC_RTOL, "max relative error %.3g" % err.max()

    announce("C4a rmat22 BC, 1024 sources with dependency rows, width %d" % width)
    xb, bb = gu.DeviceBuffer(4 * k * n), gu.DeviceBuffer(8 * n)
    try:
        bb.zero()
        st = dg.betweenness_device(src, xb, bb, cfg)
        bc = bb.to_host(np.float64)
        acc = np.zeros(n, dtype=np.float64)
        for c0 in range(0, k, 64):
            x = xb.to_host(np.float32, 64 * n, c0 * n).reshape(64, n)
            assert np.all(np.isfinite(x))
            assert np.all(x[np.arange(64), src[c0:c0 + 64]] == 0), "delta of a source is not 0"
            rs = x.sum(axis=1, dtype=np.float64)
            rel = np.abs(rs - dsum[c0:c0 + 64]) / np.maximum(dsum[c0:c0 + 64], 1.0)
            assert rel.max() <= BC_RTOL, "Brandes identity off by %.3g at source %d" % (
                rel.max(), src[c0 + int(rel.argmax())])
            acc += x.sum(axis=0, dtype=np.float64)
    finally:
        xb.close()
        bb.close()
    nz = acc != 0
    assert np.all(bc[~nz] == 0)
    assert np.max(np.abs(bc[nz] - acc[nz]) / np.abs(acc[nz])) <= 1e-9
    assert np.all(bc[np.diff(hg.R) <= 1] == 0), "BC of a leaf or isolated vertex is not 0"


# ---------------------------------------------------------------------------
# Load-balanced search, 2^26 items


def test_c4b_lbs_2p26_zipf(gu, orc, gpu):
    announce("C4b LBS 2^26 items, Zipf(1.1) counts capped at 2^20")
    c = gu.zipf_counts(1.1, 1 << 20, 1 << 26, seed=26)
    scan, res, W = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert W == oW == 1 << 26
    assert np.array_equal(scan, oscan)
    bad = np.flatnonzero(res != orc.lbs(oW, oscan))
    assert bad.size == 0, "LBS mismatch at items %s" % (bad[:5].tolist(),)


def test_c4b_lbs_2p26_100k_segments(gu, orc, gpu):
    """The same 2^26 items over 100,000 segments (the size of the reference's
    lbs.cu example), shuffled Zipf weights, with 10% of the segments (the
    first and last included) emptied and their items moved to one other."""
    announce("C4b LBS 2^26 items over 100,000 segments")
    S, W = 100000, 1 << 26
    w = 1.0 / np.arange(1, S + 1) ** 1.1
    c = np.floor(w / w.sum() * W).astype(np.int64)
    c[0] += W - c.sum()
    rng = np.random.default_rng(26)
    rng.shuffle(c)
    empty = np.concatenate([[0, S - 1], rng.choice(np.arange(1, S - 1), S // 10 - 2, replace=False)])
    moved = c[empty].sum()
    c[empty] = 0
    keep = np.flatnonzero(c)[0]
    c[keep] += moved
    assert c.sum() == W and (c == 0).sum() == S // 10
    c = c.astype(np.int32)
    scan, res, Wd = gu.lbs(c)
    oscan, oW = orc.exclusive_scan(c)
    assert Wd == oW == W
    assert np.array_equal(scan, oscan)
    assert np.array_equal(res, orc.lbs(oW, oscan))


# ---------------------------------------------------------------------------
# The query server over the serving workload's three resident graphs


def _serve_mix(gu, hgs, count, seed):
    """70% BFS (distances + path counts), 20% sampled BC, 10% diameter;
    sources per query log-uniform in [1, 512] (BC capped at 64 so that every
    BC answer can be checked against the FP64 oracle)."""
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(count):
        g = int(rng.integers(len(hgs)))
        r = rng.random()
        kind = gu.QUERY_BFS if r < 0.7 else (gu.QUERY_BC if r < 0.9 else gu.QUERY_DIAMETER)
        k = max(1, int(np.exp(rng.random() * np.log(512))))
        if kind == gu.QUERY_BC:
            k = min(k, 64)
        src = hgs[g].sample_sources(k, seed=int(rng.integers(1 << 31)))
        qs.append((g, kind, src, gu.QF_SIGMA if kind == gu.QUERY_BFS else 0))
    return qs


def _serve_run(gu, orc, hgs, dgs, qs, slots, check):
    """Runs the queries (all submitted at once when slots > 1, one at a time
    otherwise); returns per query a digest of its rows, its diameter and BC."""
    out = [None] * len(qs)
    with gu.Server(hbm_budget=48 << 30, slots=slots, max_pack=256) as srv:
        ids = [srv.add_graph(dg) for dg in dgs]

        def take(i, t, r):
            g, kind, src, _ = qs[i]
            res = srv.fetch(r, ids[g])
            srv.release(t)
            assert res["status"] == 0, "query %d status %d" % (i, res["status"])
            h = xxhash.xxh3_128()
            for key in ("dist", "sigma"):
                if key in res:
                    h.update(res[key].tobytes())
            if check and kind == gu.QUERY_BFS:
                _check_rows(orc, hgs[g], src, res["dist"], res["sigma"], "served query %d" % i)
            out[i] = (h.hexdigest(), res["diameter"], res.get("bc"), res["flags"])

        if slots == 1:
            for i, (g, kind, src, flags) in enumerate(qs):
                t = srv.submit(ids[g], kind, src, flags)
                take(i, t, srv.wait(t, 600000))
        else:
            tick = {srv.submit(ids[g], kind, src, flags): i for i, (g, kind, src, flags) in enumerate(qs)}
            for _ in qs:
                t, r = srv.next_done(600000)
                take(tick[t], t, r)
        st = srv.stats()
    assert st["completed"] == len(qs) and st["hbm_in_use"] == 0
    return out, st


def test_c5_server_resident_graphs_alone_vs_mix(gu, orc, gpu):
    announce("C5 generate R-MAT s18, R-MAT s20, 1024^2 grid")
    hgs = [gu.HostGraph.rmat(18, 16, 0.57, 0.19, 0.19, seed=18),
           gu.HostGraph.rmat(20, 16, 0.57, 0.19, 0.19, seed=20),
           gu.HostGraph.grid(1024, 0.10, seed=10)]
    dgs = [gu.DeviceGraph(h) for h in hgs]
    try:
        qs = _serve_mix(gu, hgs, 30, seed=4242)
        announce("C5 server: %d queries in a mix (4 slots)" % len(qs))
        mix, st = _serve_run(gu, orc, hgs, dgs, qs, 4, True)
        assert st["launches"] < sum((len(q[2]) + 255) // 256 for q in qs)  # queries shared launches
        announce("C5 server: the same queries one at a time")
        alone, _ = _serve_run(gu, orc, hgs, dgs, qs, 1, False)
    finally:
        for dg in dgs:
            dg.close()
    wrapped_bc = 0
    for i, (g, kind, src, _) in enumerate(qs):
        (ha, da, ba, fa), (hb, db, bb, fb) = alone[i], mix[i]
        assert ha == hb, "query %d: rows differ alone and in the mix" % i
        assert da == db
        hg = hgs[g]
        # GU_QR_SIGMA_WRAPPED: exact per source on the per-source engine (the
        # grid); R-MAT counts stay far below 2^64
        wraps = kind != gu.QUERY_DIAMETER and any(orc.sigma_wraps(hg.R, hg.C, int(s)) for s in src)
        assert fa == fb == (gu.QR_SIGMA_WRAPPED if wraps else 0), "query %d flags %d/%d" % (i, fa, fb)
        if kind == gu.QUERY_DIAMETER:
            assert db == int(orc.eccentricities_mt(hg.R, hg.C, src).max())
        elif kind == gu.QUERY_BC:
            if wraps:
                # path counts wrapped mod 2^64 (as the reference's unsigned
                # long long counts do): dependencies from them are not
                # meaningful and may be inf or NaN; alone and mixed runs
                # still agree
                wrapped_bc += 1
                assert np.allclose(ba, bb, rtol=1e-12, atol=0, equal_nan=True), "BC query %d" % i
                continue
            nz = ba != 0
            assert np.all(bb[~nz] == 0)
            if nz.any():
                assert np.max(np.abs(ba[nz] - bb[nz]) / np.abs(ba[nz])) <= 1e-12
            ref = orc.bc_f64_mt(hg.R, hg.C, src, orc.default_threads())
            pos = ref > 0
            assert np.all(bb[~pos] == 0)
            if pos.any():
                assert np.max(np.abs(bb[pos] - ref[pos]) / ref[pos]) <= BC_RTOL, "BC query %d" % i
    assert wrapped_bc < sum(1 for q in qs if q[1] == gu.QUERY_BC), "every BC query wrapped"
