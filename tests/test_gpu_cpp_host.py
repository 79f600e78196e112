"""Runs the C++ tests of the host interface above the C ABI
(tests/cpp/test_host_api.cpp, built in-tree by __graft_entry__.build()):
the reference's c2p round-trip sweep, a blast substep against the oracle, and
the interface's error behaviour.  The CPU test only checks that it was built
and links."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_host_api")


def test_cpp_host_test_built():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    r = subprocess.run(["ldd", BIN], capture_output=True, text=True)
    assert "libgrmhd_mi355x.so" in r.stdout and "not found" not in r.stdout, r.stdout


@pytest.mark.gpu
def test_gpu_cpp_host_api():
    print("[cpp host api] running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("[PASS]") == 4
