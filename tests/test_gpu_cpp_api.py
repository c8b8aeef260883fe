"""The reference's C++ skeleton tests, restated in tests/cpp/skeleton_tests.cpp
against the drop-in C++ API (skelcl_amd/include/SkelCL), run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "skelcl_cpp_tests")


def test_cpp_api_suite():
    assert os.path.exists(BIN), "build tests/cpp first (python -m skelcl_amd.build)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-6000:])
    print(r.stderr[-3000:])
    assert r.returncode == 0
    assert "0 failed" in r.stdout
