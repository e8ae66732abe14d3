"""The C++ drop-in API test program (tests/cpp/test_api.cpp) built with
AddressSanitizer + UndefinedBehaviorSanitizer against the library whose
HOST code is built the same way (sparkit_amd/_san; device code unchanged,
every report fatal): store upload and plans, the tuner, ILU(0) level
scheduling, Matrix Market parsing, RCM, the solvers' host loops and the
C++ wrappers run under the sanitizers on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_api_san")


def test_cpp_api_under_asan_and_ubsan():
    assert os.path.exists(BIN), "build tests/cpp target san first (__graft_entry__.build())"
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300, env=env)
    out = p.stdout + p.stderr
    print(out[-4000:])
    assert p.returncode == 0, out[-4000:]
    assert "0 failures" in p.stdout
    assert "AddressSanitizer" not in out and "runtime error:" not in out
