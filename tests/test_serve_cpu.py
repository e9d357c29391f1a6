"""The query server's scheduling logic on the CPU: tests/cpp/serve_test.cpp
compiles graph-utils_amd/csrc/gu_serve.hpp over an executor that computes
launches with the oracle, and checks admission, packing, overlap, failure
isolation and that every query's result is the same alone and in a mix."""
import os
import subprocess

from conftest import ROOT


def test_scheduler_alone_vs_mix_cpu():
    cpp = os.path.join(ROOT, "tests", "cpp")
    subprocess.run(["make", "-s", "-C", cpp, os.path.join(cpp, "_build", "serve_test")], check=True)
    r = subprocess.run([os.path.join(cpp, "_build", "serve_test")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
