"""The C++ API (include/turbo_cpp/lte_turbo.h) exercised from C++: the
assertions of the reference's tests/test_main.cpp, built by
turbo_cpp_amd/build.py into turbo_cpp_amd/_build/api_test."""
import os
import subprocess

import pytest

from turbo_cpp_amd import build as B


def _run(mode):
    if not os.path.exists(B.API_TEST):
        pytest.skip("api_test not built (python -m turbo_cpp_amd.build)")
    r = subprocess.run([B.API_TEST, mode], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "All tests passed" in r.stdout


def test_cpp_api_cpu_backend():
    _run("cpu")


@pytest.mark.gpu
def test_cpp_api_gpu_backends():
    _run("gpu")
