"""Runs the C++ tests of the drop-in header (tests/cpp/test_facade.cpp, the
reference's tests/00-fft.cpp and tests/01-real.cpp re-expressed) on the GPU."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_facade")


def test_cpp_facade_reference_tests():
    assert os.path.exists(BIN), "build first: __graft_entry__.build() (make -C tests/cpp)"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "all facade tests passed" in r.stdout
