"""Runs the C++ drop-in API test (tests/cpp/test_dropin.cpp), built by
__graft_entry__.build() with g++ against include/hhs/*.hpp and libhhs.so."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "test_dropin")


def test_cpp_dropin_api():
    assert os.path.exists(BIN), "build tests/cpp/test_dropin first (make -C parallel-hhs_amd)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
