"""The per-source engine's queue-fed grid: a search call with more sources
than one launch holds (256) runs as one grid of at most 256 workgroups that
take sources from a device counter (the reference's block-strided source
loop, shuffle_based.cuh:104), so no workgroup waits for another's deeper
search.  Each case is checked bit-exact against the oracle and against the
same call run as launches of 256 sources (source_queue=2), on every graph
kind that reaches a different kernel or row path:

  - degree-4 grid, path counts: LDS level-set kernel, 16-bit scratch rows
    written back by the workgroup in the caller's vertex order;
  - the same grid, distances only: distance kernel, int scratch rows;
  - degree <= 8 random graph: cluster order without the neighbour table;
  - R-MAT (hubs, no cluster order): rows searched in place;
  - directed degree-4 graph: path counts pushed along arcs;
  - one grid and 40 small ones: sources of very uneven depth;
  - a binary tree whose widest level overflows the LDS level sets, and a
    path deeper than the 16-bit distance rows: those sources are searched
    again on the distance kernel, the others come from the queue;
  - diameter sampling (eccentricities only)."""
import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

PS = 2  # GU_ENGINE_PER_SOURCE
QUEUE, CHUNKS = 1, 2


def _oracle(orc, hg, sources, sigma=True):
    d = np.empty((len(sources), hg.n), dtype=np.int32)
    s = np.empty((len(sources), hg.n), dtype=np.uint64)
    for i, src in enumerate(sources):
        d[i], s[i] = orc.sigma(hg.R, hg.C, int(src))
    return d, (s if sigma else None)


def _undirected(gu, n, edges):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    both = np.unique(np.concatenate([e, e[:, ::-1]]), axis=0)
    R = np.zeros(n + 1, dtype=np.int32)
    np.add.at(R, both[:, 0] + 1, 1)
    R = np.cumsum(R).astype(np.int32)
    return gu.HostGraph(R, both[:, 1].astype(np.int32))


def _graph(gu, kind):
    rng = np.random.default_rng(2024)
    if kind == "grid":
        return gu.HostGraph.grid(48, 0.10, seed=3)
    if kind == "deg8":
        n = 6000
        stubs = rng.permutation(np.repeat(np.arange(n), rng.integers(0, 9, n)))
        return _undirected(gu, n, stubs[: len(stubs) // 2 * 2].reshape(-1, 2))
    if kind == "rmat":
        return gu.HostGraph.rmat(12, 8, seed=41)
    if kind == "directed4":
        n = 5000
        rows = []
        for v in range(n):
            k = int(rng.integers(0, 5))
            t = rng.choice(n - 1, size=k, replace=False)
            rows.append(np.sort(np.where(t >= v, t + 1, t)))
        R = np.zeros(n + 1, dtype=np.int32)
        R[1:] = np.cumsum([len(r) for r in rows])
        return gu.HostGraph(R, np.concatenate(rows).astype(np.int32), directed=True)
    if kind == "tree":
        n = 2 ** 15 - 1  # binary tree: 16384 leaves, one level wider than the level sets
        v = np.arange(1, n)
        return _undirected(gu, n, np.stack([v, (v - 1) // 2], 1))
    if kind == "union":  # one 64^2 grid and 40 grids of 8^2: very uneven source depths
        parts = [gu.HostGraph.grid(64, 0.1, seed=5)] + [gu.HostGraph.grid(8, 0.1, seed=6 + i) for i in range(40)]
        Rs, Cs, off, base = [np.zeros(1, dtype=np.int64)], [], 0, 0
        for p in parts:
            Rs.append(p.R[1:].astype(np.int64) + base)
            Cs.append(p.C.astype(np.int64) + off)
            off += p.n
            base += p.m
        return gu.HostGraph(np.concatenate(Rs).astype(np.int32), np.concatenate(Cs).astype(np.int32))
    if kind == "path":
        n = 70000  # from an end: 69999 levels, past the 16-bit distance rows
        return _undirected(gu, n, np.stack([np.arange(n - 1), np.arange(1, n)], 1))
    raise ValueError(kind)


def _sources(hg, kind, k):
    if kind == "path":  # ends and middles: some sources rerun, some not
        return np.linspace(0, hg.n - 1, k).astype(np.int32)
    if kind == "tree":  # the root's search overflows; leaves' do not
        rng = np.random.default_rng(7)
        return np.concatenate([[0, 1, 2], rng.integers(0, hg.n, k - 3)]).astype(np.int32)
    rng = np.random.default_rng(11)
    return rng.integers(0, hg.n, k).astype(np.int32)  # repeats allowed, as in a served mix


@pytest.mark.parametrize("kind,k,sigma", [("grid", 300, True), ("grid", 517, False), ("deg8", 300, True),
                                          ("rmat", 270, True), ("directed4", 300, True),
                                          ("tree", 260, True), ("path", 260, True), ("union", 700, True)])
def test_queue_matches_oracle_and_chunks(gu, orc, gpu, kind, k, sigma):
    hg = _graph(gu, kind)
    src = _sources(hg, kind, k)
    out = {}
    with gu.DeviceGraph(hg) as dg:
        for mode in (QUEUE, CHUNKS):
            announce("source queue %s mode=%d nsrc=%d n=%d sigma=%s" % (kind, mode, k, hg.n, sigma))
            out[mode] = dg.multi_search(src, dist=True, sigma=sigma,
                                        config=gu.make_config(engine=PS, source_queue=mode))
    dq, sq, stq = out[QUEUE]
    dc, sc, stc = out[CHUNKS]
    assert stq.engine_used == PS and stc.engine_used == PS
    assert stc.batches == (k + 255) // 256
    od, os_ = _oracle(orc, hg, src, sigma)
    bad = np.argwhere(dq != od)
    assert bad.size == 0, "distance mismatch at %s" % (bad[:5].tolist(),)
    if sigma:
        bad = np.argwhere(sq != os_)
        assert bad.size == 0, "path count mismatch at %s" % (bad[:5].tolist(),)
        assert np.array_equal(sq, sc)
    assert np.array_equal(dq, dc)
    for f in ("teps_units", "reached", "max_level", "sigma_overflow"):
        assert getattr(stq, f) == getattr(stc, f), f
    if sigma:
        assert stq.dag_edges == stc.dag_edges


def test_queue_auto_above_one_launch(gu, orc, gpu):
    """source_queue=0 (the default): 256 sources run as one launch, 257 as
    the queue-fed grid (one launch, where the chunked form needs two)."""
    hg = _graph(gu, "grid")
    with gu.DeviceGraph(hg) as dg:
        for k, batches in ((256, 1), (257, 1)):
            src = _sources(hg, "grid", k)
            announce("source queue auto nsrc=%d" % k)
            d, s, st = dg.multi_search(src, True, True, gu.make_config(engine=PS))
            assert st.batches == batches
            od, os_ = _oracle(orc, hg, src[::16])
            assert np.array_equal(d[::16], od) and np.array_equal(s[::16], os_)


def test_queue_few_sources(gu, orc, gpu):
    """The queue with fewer sources than compute units (grid = sources)."""
    hg = _graph(gu, "grid")
    src = _sources(hg, "grid", 10)
    announce("source queue forced nsrc=10")
    with gu.DeviceGraph(hg) as dg:
        d, s, st = dg.multi_search(src, True, True, gu.make_config(engine=PS, source_queue=QUEUE))
    od, os_ = _oracle(orc, hg, src)
    assert np.array_equal(d, od) and np.array_equal(s, os_)


def test_queue_diameter(gu, orc, gpu):
    """Diameter sampling over 400 sources: eccentricities only, no rows."""
    hg = _graph(gu, "grid")
    src = _sources(hg, "grid", 400)
    with gu.DeviceGraph(hg) as dg:
        announce("source queue diameter nsrc=400")
        dq, stq = dg.diameter(src, gu.make_config(engine=PS, source_queue=QUEUE))
        dc, stc = dg.diameter(src, gu.make_config(engine=PS, source_queue=CHUNKS))
    expect = max(orc.eccentricity(hg.R, hg.C, int(s)) for s in np.unique(src))
    assert dq == dc == expect
    assert stq.batches == 1 and stc.batches == 2


def test_queue_rejects_bad_mode(gu, gpu):
    hg = _graph(gu, "grid")
    with gu.DeviceGraph(hg) as dg:
        with pytest.raises(gu.GuError) as ei:
            dg.multi_search(np.array([0], dtype=np.int32), config=gu.make_config(engine=PS, source_queue=3))
    assert ei.value.status == -1
