"""The C++ drop-in surface on the GPU: tests/cpp/dropin_test.cpp is a
graph-utils example driver written against the reference's signatures
(parse -> device_graph -> every *_setup -> the reference's verification),
checked with the oracle.  Built by __graft_entry__.build()."""
import os
import subprocess
import tempfile

import pytest

from conftest import ROOT, announce, graph_path

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_test")


@pytest.mark.parametrize("name", ["jazz", "test3", "test6"])
def test_dropin_driver(gpu, name):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/dropin_test is not built (run __graft_entry__.build())")
    announce("dropin_test %s" % name)
    r = subprocess.run([EXE, "-i", graph_path(name), "-f", "dimacs"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "Test failed" not in r.stdout
    assert r.stdout.count("Test passed") >= 9


@pytest.mark.parametrize("k", [700, 0])
def test_dropin_driver_more_sources_than_cus(gu, gpu, k):
    """More sources than compute units (256 on MI355X): every stored row is
    the last source s = start + r (mod 256) (multi_search_test.cu:44-63,
    common.cu:17-25), so rows wrap and the sources beyond the first 256 run
    in further launches.  A seeded 1024-vertex R-MAT written as METIS, with
    -k 700 (2 x 256 + 188 sources) and with all 1024 sources."""
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/dropin_test is not built (run __graft_entry__.build())")
    hg = gu.HostGraph.rmat(10, 16, seed=77)
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "rmat10.graph")
        hg.write_metis(path)
        announce("dropin_test rmat10 -k %d" % k)
        args = [EXE, "-i", path, "-f", "dimacs"] + (["-k", str(k)] if k else [])
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "sources=%d" % (k or hg.n) in r.stdout
    assert "Test failed" not in r.stdout
    assert r.stdout.count("Test passed") >= 9


def test_dag_reachability_dropin(gpu):
    """get_reachability_dag / reverse (graph-utils/dag_reachability.h,
    reverse.h) on random DAGs, against the oracle's BFS."""
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "reach_test")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/_build/reach_test is not built (run __graft_entry__.build())")
    announce("reach_test")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "Test passed" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
