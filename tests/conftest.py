import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
GRAPHS = os.path.join(GOLDEN, "graphs")
FIXTURES = ["test1", "test2", "test3", "test4", "test5", "test6", "jazz"]

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gu():
    import __graft_entry__

    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(gu):
    """Selects device 0; GPU tests fail loudly when none is present."""
    if gu.device_count() < 1:
        pytest.fail("no HIP device: GPU tests need an MI355X")
    return gu.choose_device(0)


def graph_path(name):
    return os.path.join(GRAPHS, name + ".graph")


def announce(msg):
    """Name the case on stdout before a GPU step runs."""
    print("[case] " + msg, flush=True)
