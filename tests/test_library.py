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


def test_blas1_entry_points_reject_bad_arguments():
    """spk_dot / spk_axpby validate their arguments before any device call:
    an unknown dtype, a negative length and null vectors with n > 0 return
    SPK_INVALID_ARGUMENT (no GPU needed)."""
    import sparkit_amd
    from sparkit_amd import _lib as L
    lib = sparkit_amd.load()
    r = ctypes.c_double()
    buf = ctypes.create_string_buffer(64)  # a non-null stand-in, never read
    vp = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.spk_dot(7, 4, vp, vp, ctypes.byref(r)) == L.INVALID_ARGUMENT
    assert b"dtype" in lib.spk_last_error()
    assert lib.spk_dot(L.F64, -1, vp, vp, ctypes.byref(r)) == L.INVALID_ARGUMENT
    assert lib.spk_dot(L.F64, 4, None, vp, ctypes.byref(r)) == L.INVALID_ARGUMENT
    assert lib.spk_dot(L.F32, 4, vp, None, ctypes.byref(r)) == L.INVALID_ARGUMENT
    assert lib.spk_dot(L.F64, 4, vp, vp, None) == L.INVALID_ARGUMENT
    assert lib.spk_axpby(3, 4, 1.0, vp, 1.0, vp) == L.INVALID_ARGUMENT
    assert b"dtype" in lib.spk_last_error()
    assert lib.spk_axpby(L.F64, -2, 1.0, vp, 1.0, vp) == L.INVALID_ARGUMENT
    assert lib.spk_axpby(L.F64, 4, 1.0, None, 1.0, vp) == L.INVALID_ARGUMENT
    assert lib.spk_axpby(L.F32, 4, 1.0, vp, 1.0, None) == L.INVALID_ARGUMENT
    assert b"null" in lib.spk_last_error()
