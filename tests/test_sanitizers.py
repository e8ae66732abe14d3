"""The oracle's C restatement under AddressSanitizer + UndefinedBehaviorSanitizer
(oracle/Makefile target `sanitize`, every report fatal): the CPU oracle tests
run against that build (SPK_ORACLE_LIB), with the sanitizer runtime preloaded
into the interpreter.  A control compiles a planted heap overflow and a
planted signed overflow with the same flags and checks that each run is
stopped with a report, so a silently inactive sanitizer cannot pass.  (The
library's host code has its own sanitized build, exercised on the GPU by
tests/test_gpu_sanitizers.py.)"""
import glob
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_LIB = os.path.join(ORACLE, "_san", "liboracle_san.so")
FLAGS = ["-O1", "-fPIC", "-shared", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"],
                       capture_output=True, text=True).stdout.strip()
    if not (os.path.isabs(p) and os.path.exists(p)):
        pytest.skip("gcc's libasan is not installed")
    return p


def _env(**extra):
    env = dict(os.environ)
    env.update(LD_PRELOAD=_libasan(),
               ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", **extra)
    return env


def test_oracle_tests_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    files = sorted(glob.glob(os.path.join(ROOT, "tests", "test_oracle_*.py")))
    assert len(files) >= 10
    p = subprocess.run(
        [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
         "-p", "no:cacheprovider", *files],
        cwd=ROOT, env=_env(SPK_ORACLE_LIB=SAN_LIB), capture_output=True,
        text=True, timeout=900)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert " passed" in p.stdout


@pytest.mark.parametrize("kind", ["heap-buffer-overflow", "signed-overflow"])
def test_sanitizer_control_catches_planted_errors(tmp_path, kind):
    body = {
        "heap-buffer-overflow": """
            #include <stdlib.h>
            double probe(int n) {
              double* a = (double*)malloc(n * sizeof(double));
              for (int i = 0; i < n; ++i) a[i] = i;
              double s = a[n];  /* one past the end */
              free(a);
              return s;
            }""",
        "signed-overflow": """
            double probe(int n) {
              int k = 2147483600;
              k += n;  /* overflows for n > 47 */
              return (double)k;
            }"""}[kind]
    src = tmp_path / "probe.c"
    src.write_text(textwrap.dedent(body))
    so = tmp_path / "probe.so"
    subprocess.run(["gcc", *FLAGS, str(src), "-o", str(so)], check=True)
    code = ("import ctypes as C; L = C.CDLL(%r); L.probe.restype = C.c_double; "
            "L.probe.argtypes = [C.c_int]; print(L.probe(64))" % str(so))
    p = subprocess.run([sys.executable, "-c", code], env=_env(),
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    expect = "AddressSanitizer" if kind == "heap-buffer-overflow" else "runtime error:"
    assert expect in p.stderr, p.stderr[-2000:]
