"""spk_cg_solve_noalloc (no reference counterpart; for callers that cannot
allocate device memory) against spk_cg_solve: the same kernels on buffers
carved from the library's static scratch array, so the same iterates, bit
for bit, on two calls in a row (the second smaller), and no device memory
taken while it runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _pair(spk, A, M, b, cfg):
    x1 = spk.DeviceVector.zeros(A.nrows, b.dtype)
    s1 = spk.conjugate_gradient(b, x1, cfg, A, M)
    x2 = spk.DeviceVector.zeros(A.nrows, b.dtype)
    s2 = spk.conjugate_gradient_noalloc(b, x2, cfg, A, M)
    return (x1, s1), (x2, s2)


def _same(a, b):
    (x1, s1), (x2, s2) = a, b
    assert s1.converged == s2.converged
    assert s1.computed_iterations == s2.computed_iterations
    assert s1.residual_norm == s2.residual_norm
    assert s1.iteration_residuals == s2.iteration_residuals
    assert np.array_equal(x1.numpy(), x2.numpy())


def test_noalloc_cg_matches_cg_twice_second_smaller(spk):
    print("no-alloc CG: poisson_3d(64), then poisson_2d(96), FP64 Jacobi",
          flush=True)
    for A in (spk.poisson_3d(64), spk.poisson_2d(96, 96, 0.25)):
        b = spk.uniform_pm1(42, A.nrows)
        cfg = spk.CGConfig(1e-8, 20000, 20000, True, spk.F64)
        r1, r2 = _pair(spk, A, spk.jacobi(A), b, cfg)
        assert r1[1].converged
        _same(r1, r2)


def test_noalloc_cg_fp32_and_explicit_inverse(spk):
    """FP32 vectors (after an ordinary FP32 solve made the FP32 copies) and
    an explicit approximate inverse in the left slot."""
    A = spk.aniso_q1_2d(128, 3, 1.5)
    M = spk.spai(spk.SpaiModel(1).predict(A).symmetric_part())
    b64 = spk.uniform_pm1(9, A.nrows)
    cfg = spk.CGConfig(1e-8, 5000, 5000, True, spk.F64)
    r1, r2 = _pair(spk, A, M, b64, cfg)
    _same(r1, r2)
    b32 = b64.astype(np.float32)
    cfg32 = spk.CGConfig(1e-5, 5000, 5000, False, spk.F32)
    J = spk.jacobi(A)
    r1, r2 = _pair(spk, A, J, b32, cfg32)  # the first call makes the copies
    _same(r1, r2)


def test_noalloc_cg_on_window_tiles(spk):
    """A window-tile store (its long rows in a store of their own): FP64, and
    FP32 once an ordinary FP32 solve has made the tiles' FP32 values."""
    print("no-alloc CG: power-law 40000 on 8x256 window tiles", flush=True)
    A = spk.powerlaw_spd(40000, 5)
    A.set_window_tiles(8, 256)
    J = spk.jacobi(A)
    b = spk.uniform_pm1(3, A.nrows)
    r1, r2 = _pair(spk, A, J, b, spk.CGConfig(1e-10, 2000, 2000, True, spk.F64))
    assert r1[1].converged
    _same(r1, r2)
    b32 = b.astype(np.float32)
    cfg32 = spk.CGConfig(1e-5, 2000, 2000, False, spk.F32)
    x32 = spk.DeviceVector.zeros(A.nrows, np.float32)
    with pytest.raises(spk.InvalidArgument, match="FP32 copy"):
        spk.conjugate_gradient_noalloc(b32, x32, cfg32, A, J)
    r1, r2 = _pair(spk, A, J, b32, cfg32)  # the first call makes the copies
    _same(r1, r2)


def test_noalloc_cg_takes_no_device_memory(spk):
    """A size no ordinary solve of this process has used: the ordinary path
    would grow its workspace; the no-allocation path leaves the device's
    free memory unchanged."""
    A = spk.poisson_3d(100, 100, 110)  # 1.1M rows: 43 MB of FP64 vectors
    J = spk.jacobi(A)
    b = spk.uniform_pm1(5, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    # a small solve first, so code loading is not counted below
    S = spk.poisson_2d(16, 16, 0.5)
    spk.conjugate_gradient_noalloc(spk.uniform_pm1(1, S.nrows),
                                   spk.DeviceVector.zeros(S.nrows),
                                   spk.CGConfig(1e-6, 100, 100), S, spk.jacobi(S))
    spk.synchronize()
    before = _hip_free_bytes()
    s = spk.conjugate_gradient_noalloc(b, x, spk.CGConfig(1e-6, 2000, 2000),
                                       A, J)
    spk.synchronize()
    after = _hip_free_bytes()
    assert s.converged
    assert abs(before - after) < (1 << 20), (before, after)
    assert spk.true_relres(A, b, x) < 1e-6


def test_noalloc_cg_rejects_what_it_cannot_hold(spk):
    A = spk.poisson_3d(128)  # 2.1M rows: more than the 64 MiB scratch holds
    b = spk.uniform_pm1(5, A.nrows)
    x = spk.DeviceVector.from_numpy(np.full(A.nrows, 3.0))
    with pytest.raises(spk.InvalidArgument, match="scratch"):
        spk.conjugate_gradient_noalloc(b, x, spk.CGConfig(1e-8, 100, 100), A,
                                       spk.jacobi(A))
    assert np.all(x.numpy() == 3.0)  # nothing was queued
    S = spk.poisson_2d(40)
    bs = spk.uniform_pm1(5, S.nrows)
    xs = spk.DeviceVector.zeros(S.nrows)
    with pytest.raises(spk.InvalidArgument, match="ILU"):
        spk.conjugate_gradient_noalloc(bs, xs, spk.CGConfig(1e-8, 100, 100), S,
                                       spk.ilu0(S))
    fresh = spk.poisson_2d(24, 24, 0.5)  # no FP32 copy made yet
    b32 = spk.DeviceVector.zeros(fresh.nrows, np.float32)
    x32 = spk.DeviceVector.zeros(fresh.nrows, np.float32)
    with pytest.raises(spk.InvalidArgument, match="FP32 copy"):
        spk.conjugate_gradient_noalloc(
            b32, x32, spk.CGConfig(1e-6, 100, 100, False, spk.F32), fresh)
