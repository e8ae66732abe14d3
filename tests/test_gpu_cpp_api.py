"""Runs the C++ drop-in API test program (tests/cpp/test_api.cpp, built by
__graft_entry__.build()) on the GPU: the reference's solver signatures with
device operator/preconditioner objects."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_api")


def test_cpp_api():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build())"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout
