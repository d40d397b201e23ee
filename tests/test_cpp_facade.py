"""Runs the C++ restatement of tests/correctness.rs (tests/cpp/correctness.cpp)
over the C++ facade (flash-map_amd/host/flashmap.hpp) on each backend."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "flash-map_amd", "_build", "correctness")


def _run(backend):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (make -C flash-map_amd)")
    r = subprocess.run([BIN, backend], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "18 tests, 0 failed" in r.stdout


def test_cpp_reference_tests_cpu_backend():
    _run("cpu")


@pytest.mark.gpu
def test_cpp_reference_tests_gpu_backend():
    _run("gpu")
