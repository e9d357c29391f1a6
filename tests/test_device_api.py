"""GPU tests of the public device API, include/graph_utils_amd_device.hpp.

The kernels under test are in tests/cpp/device_api_test.hip, written against
that header only, as a caller's kernels would be.  Contracts checked:

- race_and_resolve_wave (race_and_resolve.cuh:13-25): every lane gets the
  highest lane whose predicate holds, or its own lane id when none does;
  checked over every predicate pattern of 16 lanes (at both ends of the
  wave), every single and pair of lanes, and random 64-lane masks;
- atomic_agg_inc (shuffle_based.cuh:58-76): the participating lanes take
  consecutive slots in lane order, all slots of a launch are distinct and
  the counter ends at the number of participants (global and LDS counters,
  branch and predicate forms);
- get_bit / set_bit (kernel_utils.cuh:22-35) against Python bit arithmetic;
- extract_edges_block / extract_edges_wave (load_balanced_search.cu:63-202):
  scan and search bit-exact against the oracle (lbs.cu:18-31), at the
  reference's 512-thread, 2-items-per-thread shape and others, with
  zero-count segments and tile-boundary sizes; several wave searches in one
  workgroup do not interfere;
- multi_search<F1..F7> (shuffle_based.cuh:80-286) with hook sets for
  distances, path counts, diameter sampling and Brandes: each stored row
  equals the oracle's row for the source the reference's [start, end) row
  mapping puts there (multi_search_test.cu:44-63).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, announce, graph_path

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "tests", "cpp", "_build", "libdevapi_test.so")
P = ctypes.c_void_p
i32, i64 = ctypes.c_int32, ctypes.c_int64
INT_MAX = 2147483647


@pytest.fixture(scope="module")
def dt(gpu):
    if not os.path.exists(LIB):
        pytest.fail("tests/cpp/_build/libdevapi_test.so is not built (make -C tests/cpp)")
    L = ctypes.CDLL(LIB)
    L.dt_race.argtypes = [P, i32, P, P]
    L.dt_agg_inc.argtypes = [P, i32, i32, P, P, P]
    L.dt_bits.argtypes = [P, P, P, i32, P, P]
    L.dt_lbs.argtypes = [P, i32, i32, i32, P, P, i64, P]
    L.dt_lbs_waves.argtypes = [P, i64, P, P, P, i32, i64, P, P, P]
    L.dt_multi_search.argtypes = [P, P, i32, i64, i32, i32, i32, i32, i32, P, P, P, P]
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _masks():
    rng = np.random.default_rng(7)
    low = np.arange(1 << 16, dtype=np.uint64)
    high = low << np.uint64(48)
    single = np.uint64(1) << np.arange(64, dtype=np.uint64)
    pairs = np.array([(1 << a) | (1 << b) for a in range(64) for b in range(a + 1, 64)], dtype=np.uint64)
    rand = rng.integers(0, 1 << 63, 4096, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, 4096).astype(np.uint64)
    ends = np.array([0, (1 << 64) - 1, 1 << 63, 1], dtype=np.uint64)
    return np.concatenate([low, high, single, pairs, rand, ends])


def test_race_and_resolve_wave(dt):
    m = _masks()
    announce("race_and_resolve_wave over %d masks" % len(m))
    out = np.empty((len(m), 64), dtype=np.int32)
    lanes = np.empty((len(m), 64), dtype=np.int32)
    assert dt.dt_race(_p(m), len(m), _p(out), _p(lanes)) == 0
    exp = np.empty((len(m), 64), dtype=np.int32)
    lane = np.arange(64, dtype=np.int32)
    for i, x in enumerate(m.tolist()):
        exp[i] = (x.bit_length() - 1) if x else lane
    bad = np.argwhere(out != exp)
    assert bad.size == 0, "mask %#x lane %d: got %d" % (int(m[bad[0][0]]), bad[0][1], out[tuple(bad[0])])
    # lane_id / wave_id: four waves per workgroup
    assert np.array_equal(lanes, lane[None, :] + 64 * (np.arange(len(m)) % 4)[:, None])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_atomic_agg_inc(dt, mode):
    rng = np.random.default_rng(mode)
    m = np.concatenate([np.array([0, (1 << 64) - 1, 1, 1 << 63], dtype=np.uint64),
                        rng.integers(0, 1 << 63, 1020, dtype=np.int64).astype(np.uint64) << np.uint64(1)])
    announce("atomic_agg_inc mode %d" % mode)
    slots = np.empty((len(m), 64), dtype=np.int32)
    total = np.zeros(1, dtype=np.int32)
    bt = np.zeros((len(m) + 3) // 4, dtype=np.int32)
    assert dt.dt_agg_inc(_p(m), len(m), mode, _p(slots), _p(total), _p(bt)) == 0
    part = ((m[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
    assert np.all(slots[~part] == -1)
    for i in range(len(m)):
        s = slots[i][part[i]]
        if len(s):  # consecutive, in lane order
            assert np.array_equal(s, s[0] + np.arange(len(s))), "mask %d" % i
    if mode == 2:
        for b in range(len(bt)):
            s = slots[4 * b:4 * b + 4][part[4 * b:4 * b + 4]]
            assert bt[b] == len(s) and np.array_equal(np.sort(s), np.arange(len(s)))
    else:
        s = np.sort(slots[part])
        assert total[0] == part.sum() and np.array_equal(s, np.arange(part.sum()))


def test_get_set_bit(dt):
    rng = np.random.default_rng(3)
    n = 1 << 16
    vals = rng.integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32)
    pos = rng.integers(0, 32, n).astype(np.int32)
    ins = rng.integers(0, 4, n).astype(np.uint32)
    got = np.empty(n, dtype=np.int32)
    st = np.empty(n, dtype=np.uint32)
    announce("get_bit / set_bit")
    assert dt.dt_bits(_p(vals), _p(pos), _p(ins), n, _p(got), _p(st)) == 0
    u = vals.view(np.uint32).astype(np.uint64)
    one = np.uint64(1)
    assert np.array_equal(got, ((u >> pos.astype(np.uint64)) & one).astype(np.int32))
    exp = (u & ~(one << pos.astype(np.uint64))) | ((ins.astype(np.uint64) & one) << pos.astype(np.uint64))
    assert np.array_equal(st, (exp & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def _lbs_cases():
    rng = np.random.default_rng(11)
    # the reference example's segment count with counts it can hold in int32
    ref = rng.integers(0, 100, 100000).astype(np.int32)
    ref[:3] = 0
    ref[50000:50010] = 0
    ref[-4:] = 0
    cases = {"ref_100k": ref}
    for S in (1, 63, 64, 65, 1023, 1024, 1025, 2049):
        c = rng.integers(0, 40, S).astype(np.int32)
        c[rng.random(S) < 0.2] = 0
        cases["S%d" % S] = c
    cases["one_zero"] = np.zeros(1, np.int32)
    cases["all_zero"] = np.zeros(3000, np.int32)
    cases["one_big"] = np.array([0, 0, 300000, 0], np.int32)
    return cases


def _run_lbs(dt, orc, c, form, block):
    oscan, oW = orc.exclusive_scan(c)
    scan = np.empty(len(c), dtype=np.int32)
    res = np.empty(max(oW, 1), dtype=np.int32)
    W = np.zeros(1, dtype=np.int32)
    assert dt.dt_lbs(_p(c), len(c), form, block, _p(scan), _p(res), max(oW, 1), _p(W)) == 0
    assert W[0] == oW
    assert np.array_equal(scan, oscan)
    if oW:
        exp = orc.lbs(oW, oscan)
        bad = np.flatnonzero(res[:oW] != exp)
        assert bad.size == 0, "item %d: got %d, want %d" % (bad[0], res[bad[0]], exp[bad[0]])


@pytest.mark.parametrize("block", [512, 64, 256, 1024])
def test_extract_edges_block(dt, orc, block):
    for name, c in _lbs_cases().items():
        announce("extract_edges_block<<<1,%d>>> %s" % (block, name))
        _run_lbs(dt, orc, c, 1, block)


def test_extract_edges_wave(dt, orc):
    for name, c in _lbs_cases().items():
        if len(c) > 5000:
            c = c[:5000].copy()
        announce("extract_edges_wave<<<1,64>>> %s" % name)
        _run_lbs(dt, orc, c, 0, 64)


def test_wave_searches_do_not_interfere(dt, orc):
    rng = np.random.default_rng(5)
    lens = [1, 100, 777, 3000, 64, 129, 5, 2000]
    counts = [rng.integers(0, 30, L).astype(np.int32) for L in lens]
    off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int32)
    Ws = [int(c.astype(np.int64).sum()) for c in counts]
    roff = np.concatenate([[0], np.cumsum(Ws)[:-1]]).astype(np.int64)
    allc = np.concatenate(counts)
    scan = np.empty(len(allc), dtype=np.int32)
    res = np.empty(max(sum(Ws), 1), dtype=np.int32)
    tot = np.empty(len(lens), dtype=np.int32)
    announce("8 wave searches in one workgroup")
    assert dt.dt_lbs_waves(_p(allc), len(allc), _p(off), _p(np.array(lens, np.int32)), _p(roff), len(lens),
                           max(sum(Ws), 1), _p(scan), _p(res), _p(tot)) == 0
    for j, c in enumerate(counts):
        oscan, oW = orc.exclusive_scan(c)
        assert tot[j] == oW
        assert np.array_equal(scan[off[j]:off[j] + lens[j]], oscan)
        assert np.array_equal(res[roff[j]:roff[j] + oW], orc.lbs(oW, oscan))


def _graphs(gu):
    return {"jazz": gu.HostGraph.parse_metis(graph_path("jazz")),
            "test6": gu.HostGraph.parse_metis(graph_path("test6")),
            "rmat12": gu.HostGraph.rmat(12, 16, seed=31)}


@pytest.mark.parametrize("kind", ["bfs", "sigma", "diameter", "bc"])
def test_multi_search_hooks(gu, orc, dt, kind):
    k = {"bfs": 0, "sigma": 1, "diameter": 2, "bc": 3}[kind]
    for name, hg in _graphs(gu).items():
        n = hg.n
        for (start, end, G, block) in ((0, n, 16, 1024), (1, min(n, 300), 256, 256), (2, min(n, 9), 4, 64)):
            announce("multi_search hooks %s on %s [%d,%d) G=%d block=%d" % (kind, name, start, end, G, block))
            d = np.empty((G, n), dtype=np.int32)
            s = np.empty((G, n), dtype=np.uint64)
            x = np.empty((G, n), dtype=np.float32)
            diam = np.zeros(1, dtype=np.int32)
            assert dt.dt_multi_search(_p(hg.R), _p(hg.C), n, hg.m, start, end, k, G, block, _p(d), _p(s),
                                      _p(x), _p(diam)) == 0
            rows = orc.rows_for_range(start, end, G)
            if kind == "diameter":
                assert diam[0] == max(orc.eccentricity(hg.R, hg.C, i) for i in range(start, end))
            for r, src in enumerate(rows):
                od, os_ = orc.sigma(hg.R, hg.C, int(src))
                assert np.array_equal(d[r], od), "%s row %d (source %d): distances" % (name, r, src)
                if kind in ("sigma", "bc"):
                    assert np.array_equal(s[r], os_), "%s row %d (source %d): path counts" % (name, r, src)
                if kind == "bc":
                    ref = orc.brandes_f64(hg.R, hg.C, int(src))
                    pos = ref > 0
                    assert np.all(x[r][~pos] == 0)
                    if pos.any():
                        err = np.abs(x[r][pos] - ref[pos]) / ref[pos]
                        assert err.max() <= 1e-5, "%s source %d: delta error %.3g" % (name, src, err.max())
