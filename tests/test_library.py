"""CPU checks of the C ABI library: it loads, exports every entry point the
public header declares, and fails cleanly without a GPU (no compute calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sparkit_hip.h")
LIB = os.path.join(ROOT, "sparkit_amd", "libsparkit_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(spk_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"


def test_every_declared_symbol_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (spk_[a-z0-9_]+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_python_binding_covers_header():
    import sparkit_amd._lib as L
    assert sorted(L.exported_names()) == declared()


def test_loads_and_reports_status_without_gpu():
    import sparkit_amd
    lib = sparkit_amd.load()
    have_gpu = os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK)
    if have_gpu:
        pytest.skip("a GPU is present; covered by the gpu tests")
    st = lib.spk_init(0)
    assert st == sparkit_amd._lib.NO_DEVICE
    assert b"device" in lib.spk_last_error()


def test_gfx950_code_object():
    """The library carries gfx950 device code."""
    data = open(LIB, "rb").read()
    assert b"gfx950" in data
