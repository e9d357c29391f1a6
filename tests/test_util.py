"""util_device beyond timers: the power sampler that replaces the
reference's NVML power_measurer (util_device.cu:79-130) and the roctx trace
ranges.  The CPU tests check argument handling and that the ranges are
harmless without a profiler; the GPU tests sample real power around a
search, through the C ABI and through the C++ drop-in driver's -n flag."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, announce, graph_path

EXE = os.path.join(ROOT, "tests", "cpp", "_build", "dropin_test")


def test_trace_ranges_nest_without_a_profiler(gu):
    L = gu.lib()
    first = L.gu_trace_push(b"outer")
    assert first in (0, gu.ERR_UNSUPPORTED)
    assert L.gu_trace_push(b"inner") == first
    assert L.gu_trace_pop() == first
    assert L.gu_trace_pop() == first
    assert L.gu_trace_push(None) == -1  # GU_ERR_INVALID
    with gu.TraceRange("python range") as r:
        assert r.on == (first == 0)


def test_power_start_rejects_bad_arguments(gu):
    L = gu.lib()
    h = ctypes.c_void_p()
    assert L.gu_power_start(0, 10, None) == -1
    assert L.gu_power_start(0, 0, ctypes.byref(h)) == -1
    assert L.gu_power_start(0, -5, ctypes.byref(h)) == -1
    assert L.gu_power_stop(None, None) == -1


@pytest.mark.gpu
def test_power_sampling_around_a_search(gu, gpu):
    """Mean socket power over a ~1 s stretch of searches: either the host
    reads the device's power (samples taken, plausible watts, the energy
    counter's average beside them) or the sampler says it cannot
    (end() == -1, the reference's answer on boards NVML cannot read)."""
    announce("power sampling around rmat14 searches")
    hg = gu.HostGraph.rmat(14, 16, seed=3)
    src = hg.sample_sources(64, seed=4)
    pm = gu.PowerMeasurer(10)
    with gu.DeviceGraph(hg) as dg:
        started = pm.start()
        with gu.Timer() as t:
            reps = 0
            while reps < 400:
                dg.multi_search(src, dist=True, sigma=True)
                reps += 1
                if reps % 20 == 0:
                    gu.synchronize()
        avg = pm.end()
    print("power: started=%s status=%s avg=%.1f W over %.3f s of searches" % (started, pm.status, avg, t.ms / 1e3))
    if not started:
        assert pm.status == gu.ERR_UNSUPPORTED and avg == -1.0
        return
    r = pm.result
    print("samples=%d max=%.1f W energy_avg=%.1f W energy=%.2f J seconds=%.3f"
          % (r.samples, r.max_w, r.energy_avg_w, r.energy_j, r.seconds))
    assert r.samples >= 1
    assert 0.0 < avg <= r.max_w < 5000.0
    assert r.seconds > 0
    if r.energy_avg_w >= 0:
        assert 0.0 < r.energy_avg_w < 5000.0


@pytest.mark.gpu
def test_dropin_driver_power_flag(gpu):
    """-n/--nvml: the drop-in driver prints the average power around its
    first search (multi_search_test.cu:173-179) and still passes."""
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_build/dropin_test is not built (run __graft_entry__.build())")
    announce("dropin_test jazz -n")
    r = subprocess.run([EXE, "-i", graph_path("jazz"), "-f", "dimacs", "-n"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-1000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "Average power:" in r.stdout
    line = [x for x in r.stdout.splitlines() if x.startswith("Average power:")][0]
    w = float(line.split()[2])
    assert w == -1.0 or 0.0 < w < 5000.0
