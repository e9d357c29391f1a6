"""GPU tests of the per-source output forms (gu_multi_search_rows,
gu_betweenness_rows), shared graph handles and the query server: every
query's result is the same alone and in a mix of concurrent, packed
launches, and equals the oracle (distances / path counts bit-exact, BC
within 1e-5 relative of FP64)."""
import threading

import numpy as np
import pytest

from conftest import announce, graph_path

pytestmark = pytest.mark.gpu


def _oracle_rows(orc, hg, sources):
    d = np.stack([orc.sigma(hg.R, hg.C, int(s))[0] for s in sources])
    s = np.stack([orc.sigma(hg.R, hg.C, int(s))[1] for s in sources])
    return d, s


def _ecc(orc, hg, sources):
    return np.array([orc.eccentricity(hg.R, hg.C, int(s)) for s in sources], dtype=np.int32)


@pytest.mark.parametrize("graph,engine", [("rmat12", 1), ("grid40", 2), ("jazz", 1), ("jazz", 2)])
def test_rows_match_contiguous_and_oracle(gu, orc, gpu, graph, engine):
    announce("rows %s engine %d" % (graph, engine))
    if graph == "rmat12":
        hg = gu.HostGraph.rmat(12, 16, seed=12)
    elif graph == "grid40":
        hg = gu.HostGraph.grid(40, 0.1, seed=40)
    else:
        hg = gu.HostGraph.parse_metis(graph_path("jazz"))
    n = hg.n
    src = hg.sample_sources(70, seed=5)
    cfg = gu.make_config(engine=engine)
    with gu.DeviceGraph(hg) as dg:
        d_ref, s_ref, _ = dg.multi_search(src, dist=True, sigma=True, config=cfg)
        # one buffer per source, rows handed over in scrambled order
        perm = np.random.default_rng(1).permutation(len(src))
        dbufs = [gu.DeviceBuffer(4 * n) for _ in src]
        sbufs = [gu.DeviceBuffer(8 * n) for _ in src]
        try:
            ecc, st = dg.multi_search_rows(src[perm], [dbufs[i].ptr.value for i in perm],
                                           [sbufs[i].ptr.value for i in perm], want_ecc=True,
                                           config=cfg)
            assert st.engine_used == engine
            for j, i in enumerate(perm):
                assert np.array_equal(dbufs[i].to_host(np.int32), d_ref[i])
                assert np.array_equal(sbufs[i].to_host(np.uint64), s_ref[i])
            assert np.array_equal(ecc, _ecc(orc, hg, src[perm]))
            # eccentricities alone (no rows)
            ecc2, _ = dg.multi_search_rows(src, None, None, want_ecc=True, config=cfg)
            assert np.array_equal(ecc2, _ecc(orc, hg, src))
        finally:
            for b in dbufs + sbufs:
                b.close()
    d_o, s_o = _oracle_rows(orc, hg, src[:8])
    assert np.array_equal(d_ref[:8], d_o) and np.array_equal(s_ref[:8], s_o)


def test_betweenness_rows_groups(gu, orc, gpu):
    announce("betweenness_rows groups rmat12")
    hg = gu.HostGraph.rmat(12, 16, seed=13)
    n = hg.n
    src = hg.sample_sources(100, seed=6)
    a, b = src[:37], src[37:]
    with gu.DeviceGraph(hg) as dg:
        _, bc_a, _ = dg.betweenness(a, delta=False)
        _, bc_b, _ = dg.betweenness(b, delta=False)
        xa, _, _ = dg.betweenness(a, delta=True, bc=False)
        ta, tb = gu.DeviceBuffer(8 * n), gu.DeviceBuffer(8 * n)
        xb = [gu.DeviceBuffer(4 * n) for _ in range(len(a))]
        try:
            ta.zero()
            tb.zero()
            targets = [ta.ptr.value] * len(a) + [tb.ptr.value] * len(b)
            rows = [x.ptr.value for x in xb] + [0] * len(b)
            dg.betweenness_rows(src, rows, targets)
            got_a, got_b = ta.to_host(np.float64), tb.to_host(np.float64)
            for i in range(len(a)):
                assert np.array_equal(xb[i].to_host(np.float32), xa[i])
        finally:
            for x in [ta, tb] + xb:
                x.close()
    for got, ref in ((got_a, bc_a), (got_b, bc_b)):
        nz = ref != 0
        assert np.all(got[~nz] == 0)
        assert np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])) <= 1e-12
    ref64 = orc.bc_f64(hg.R, hg.C, a)
    nz = ref64 > 0
    assert np.max(np.abs(got_a[nz] - ref64[nz]) / ref64[nz]) <= 1e-5


def test_shared_handles_run_concurrently(gu, orc, gpu):
    announce("shared handles, two threads")
    hg = gu.HostGraph.rmat(13, 16, seed=14)
    srcs = [hg.sample_sources(64, seed=s) for s in (1, 2, 3, 4)]
    with gu.DeviceGraph(hg) as dg:
        ref = [dg.multi_search(s, dist=True, sigma=True)[:2] for s in srcs]
        views = [dg.share(), dg.share()]
        got = [None] * 4
        errs = []

        def work(t):
            try:
                for i in range(t, 4, 2):
                    got[i] = views[t].multi_search(srcs[i], dist=True, sigma=True)[:2]
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert gu.lib().gu_graph_destroy(dg.h) == -1  # refused while shared handles exist
        for v in views:
            v.close()
    assert not errs, errs
    for i in range(4):
        assert np.array_equal(got[i][0], ref[i][0]) and np.array_equal(got[i][1], ref[i][1])


def _make_queries(gu, graphs, count, seed, max_src):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(count):
        g = int(rng.integers(len(graphs)))
        r = rng.random()
        kind = gu.QUERY_BFS if r < 0.6 else (gu.QUERY_BC if r < 0.85 else gu.QUERY_DIAMETER)
        k = max(1, int(np.exp(rng.random() * np.log(max_src))))
        src = graphs[g].sample_sources(k, seed=int(rng.integers(1 << 30)))
        flags = gu.QF_SIGMA if (kind == gu.QUERY_BFS and rng.random() < 0.8) else 0
        if kind == gu.QUERY_BC and rng.random() < 0.3:
            flags = gu.QF_DELTA
        qs.append((g, kind, src, flags))
    return qs


def _run(gu, dgs, qs, budget, slots, max_pack, sequential):
    out = [None] * len(qs)
    with gu.Server(hbm_budget=budget, slots=slots, max_pack=max_pack) as srv:
        ids = [srv.add_graph(dg) for dg in dgs]
        if sequential:
            for i, (g, kind, src, flags) in enumerate(qs):
                t = srv.submit(ids[g], kind, src, flags)
                r = srv.wait(t)
                out[i] = srv.fetch(r, ids[g])
                out[i]["launches"] = r.launches
                srv.release(t)
        else:
            tick = {}
            for i, (g, kind, src, flags) in enumerate(qs):
                tick[srv.submit(ids[g], kind, src, flags)] = i
            for _ in qs:
                t, r = srv.next_done(60000)
                i = tick[t]
                out[i] = srv.fetch(r, ids[qs[i][0]])
                out[i]["launches"] = r.launches
                srv.release(t)
        st = srv.stats()
    return out, st


def test_server_alone_vs_mix(gu, orc, gpu):
    announce("server alone vs mix")
    graphs = [gu.HostGraph.rmat(12, 16, seed=21), gu.HostGraph.grid(48, 0.1, seed=22),
              gu.HostGraph.parse_metis(graph_path("jazz"))]
    dgs = [gu.DeviceGraph(h) for h in graphs]
    try:
        qs = _make_queries(gu, graphs, 48, seed=3, max_src=80)
        big = max(len(q[2]) for q in qs) * max(h.n for h in graphs) * 12
        alone, _ = _run(gu, dgs, qs, 2 * big, 1, 256, True)
        mix, st = _run(gu, dgs, qs, 3 * big, 4, 64, False)
    finally:
        for dg in dgs:
            dg.close()
    assert st["completed"] == len(qs) and st["hbm_in_use"] == 0
    assert st["launches"] < sum((len(q[2]) + 63) // 64 for q in qs)  # queries shared launches
    for i, (g, kind, src, flags) in enumerate(qs):
        a, b = alone[i], mix[i]
        assert a["status"] == 0 and b["status"] == 0
        for key in ("dist", "sigma", "delta"):
            assert (key in a) == (key in b)
            if key in a:
                assert np.array_equal(a[key], b[key]), (i, key)
        assert a["diameter"] == b["diameter"]
        hg = graphs[g]
        if kind == gu.QUERY_BFS:
            d_o, s_o = _oracle_rows(orc, hg, src[:4])
            assert np.array_equal(b["dist"][:4], d_o)
            if "sigma" in b:
                assert np.array_equal(b["sigma"][:4], s_o)
        elif kind == gu.QUERY_DIAMETER:
            assert b["diameter"] == int(_ecc(orc, hg, src).max())
        else:
            nz = a["bc"] != 0
            assert np.all(b["bc"][~nz] == 0)
            if nz.any():
                assert np.max(np.abs(a["bc"][nz] - b["bc"][nz]) / np.abs(a["bc"][nz])) <= 1e-12
            ref = orc.bc_f64(hg.R, hg.C, src)
            pos = ref > 0
            assert np.all(b["bc"][~pos] == 0)
            if pos.any():
                assert np.max(np.abs(b["bc"][pos] - ref[pos]) / ref[pos]) <= 1e-5


def test_server_host_delivery(gu, orc, gpu):
    """GU_QF_HOST: outputs copied to the server's pinned host arena before a
    query is reported done; the same answers as device-resident results,
    device pointers NULL and the device block given back at completion."""
    announce("server host delivery")
    graphs = [gu.HostGraph.rmat(12, 16, seed=51), gu.HostGraph.grid(48, 0.1, seed=52),
              gu.HostGraph.parse_metis(graph_path("jazz"))]
    dgs = [gu.DeviceGraph(h) for h in graphs]
    try:
        qs = _make_queries(gu, graphs, 40, seed=9, max_src=80)
        big = max(len(q[2]) for q in qs) * max(h.n for h in graphs) * 12
        dev, _ = _run(gu, dgs, qs, 3 * big, 4, 64, False)
        hq = [(g, k, src, fl | gu.QF_HOST) for (g, k, src, fl) in qs]
        out = [None] * len(hq)
        with gu.Server(hbm_budget=3 * big, slots=4, max_pack=64, host_budget=4 * big) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            tick = {srv.submit(ids[g], k, src, fl): i for i, (g, k, src, fl) in enumerate(hq)}
            for _ in hq:
                t, r = srv.next_done(60000)
                i = tick[t]
                assert r.status == 0
                assert not (r.dist_dev or r.sigma_dev or r.delta_dev or r.bc_dev)
                out[i] = srv.fetch(r, ids[hq[i][0]])
                srv.release(t)
            st = srv.stats()
        assert st["completed"] == len(hq) and st["hbm_in_use"] == 0 and st["host_in_use"] == 0
        assert 0 < st["host_peak"] <= 4 * big
        with gu.Server(hbm_budget=big, slots=1) as srv:  # no host arena: refused
            gid = srv.add_graph(dgs[2])
            with pytest.raises(gu.GuError) as e:
                srv.submit(gid, gu.QUERY_BFS, np.array([1], np.int32), gu.QF_HOST)
            assert e.value.status == gu.ERR_INVALID
    finally:
        for dg in dgs:
            dg.close()
    for i, (g, kind, src, flags) in enumerate(qs):
        a, b = dev[i], out[i]
        for key in ("dist", "sigma", "delta"):
            assert (key in a) == (key in b), (i, key)
            if key in a:
                assert np.array_equal(a[key], b[key]), (i, key)
        assert a["diameter"] == b["diameter"] and a["flags"] == b["flags"]
        if kind == gu.QUERY_BC:
            nz = a["bc"] != 0
            assert np.all(b["bc"][~nz] == 0)
            if nz.any():
                assert np.max(np.abs(a["bc"][nz] - b["bc"][nz]) / np.abs(a["bc"][nz])) <= 1e-12
        if kind == gu.QUERY_BFS:
            d_o, s_o = _oracle_rows(orc, graphs[g], src[:3])
            assert np.array_equal(b["dist"][:3], d_o)
            if "sigma" in b:
                assert np.array_equal(b["sigma"][:3], s_o)


def test_server_rejects_oversized_query(gu, gpu):
    announce("server oversized query")
    hg = gu.HostGraph.rmat(10, 8, seed=2)
    with gu.DeviceGraph(hg) as dg, gu.Server(hbm_budget=1 << 20, slots=2) as srv:
        gid = srv.add_graph(dg)
        t = srv.submit(gid, gu.QUERY_BFS, hg.sample_sources(200, seed=1), gu.QF_SIGMA)
        r = srv.wait(t, 10000)
        assert r.status == gu.ERR_CAPACITY
        t2 = srv.submit(gid, gu.QUERY_DIAMETER, hg.sample_sources(5, seed=2))
        r2 = srv.wait(t2, 10000)
        assert r2.status == 0 and r2.diameter > 0
        srv.release(t)
        srv.release(t2)
        assert srv.stats()["rejected"] == 1


def test_server_queries_allocate_nothing(gu, gpu):
    """Registering a graph reserves each slot's search workspace: a stream of
    every query kind at 1, 17 and 256 sources, on a graph of the batched
    engine and one of the per-source level-set engine, then leaves the
    device's free memory where it was (within 16 MiB for the runtime)."""
    announce("server reserves workspace")
    hgs = [gu.HostGraph.rmat(16, 16, seed=16), gu.HostGraph.grid(512, 0.1, seed=51)]
    dgs = [gu.DeviceGraph(h) for h in hgs]
    try:
        for dg in dgs:  # loads the kernels; sizes the handles' own workspaces
            dg.diameter(np.arange(4, dtype=np.int32))
        with gu.Server(hbm_budget=2 << 30, slots=2) as srv:
            ids = [srv.add_graph(dg) for dg in dgs]
            free0 = gu.choose_device(0)["free_mem"]
            ticks = []
            for g, hg in enumerate(hgs):
                for k in (256, 17, 1):
                    src = hg.sample_sources(k, seed=k)
                    for kind, fl in ((gu.QUERY_BFS, gu.QF_SIGMA), (gu.QUERY_BFS, 0), (gu.QUERY_BC, 0),
                                     (gu.QUERY_DIAMETER, 0)):
                        ticks.append(srv.submit(ids[g], kind, src, fl))
            for t in ticks:
                assert srv.wait(t, 60000).status == 0
                srv.release(t)
            free1 = gu.choose_device(0)["free_mem"]
    finally:
        for dg in dgs:
            dg.close()
    assert free0 - free1 <= 16 << 20, (free0, free1)


def test_autotune_keeps_results_exact(gu, orc, gpu):
    announce("autotune rmat12 / grid32")
    hg = gu.HostGraph.rmat(12, 16, seed=41)
    src = hg.sample_sources(128, seed=42)
    with gu.DeviceGraph(hg) as dg:
        ref_d, ref_s, _ = dg.multi_search(src, True, True)
        t = dg.autotune(src, gu.QUERY_BFS)
        assert t["trials"] >= 10 and t["ms_tuned"] <= t["ms_default"]
        d, s, st = dg.multi_search(src, True, True)
        assert st.engine_used == t["engine"]
        if t["engine"] == 1:
            assert st.variant_used == t["variant"]
        assert np.array_equal(d, ref_d) and np.array_equal(s, ref_s)
        tb = dg.autotune(src[:64], gu.QUERY_BC)
        _, bc, _ = dg.betweenness(src[:64], delta=False)
        view = dg.share()
        _, bc2, st2 = view.betweenness(src[:64], delta=False)
        assert st2.engine_used == tb["engine"]
        view.close()
    ref = orc.bc_f64(hg.R, hg.C, src[:64])
    nz = ref > 0
    assert np.max(np.abs(bc[nz] - ref[nz]) / ref[nz]) <= 1e-5
    assert np.array_equal(bc, bc2)
    hg = gu.HostGraph.grid(32, 0.1, seed=43)
    with gu.DeviceGraph(hg) as dg:
        t = dg.autotune(hg.sample_sources(64, seed=44), gu.QUERY_BFS)
        assert t["engine"] in (1, 2)
