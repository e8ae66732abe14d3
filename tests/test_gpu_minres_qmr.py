"""GPU parity of the device MINRES (minres.hpp) and QMR (qmr.hpp) against the
oracle: iterates stopped at any step, convergence, iteration counts,
residual-estimate histories, the preconditioner slots (diagonal ones fused
into the vector passes, ILU(0) through the general apply path), FP32 mode
and the edge cases of the reference's tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dv(spk, a, dt=np.float64):
    return spk.DeviceVector.from_numpy(np.asarray(a, dtype=dt))


def _prec(spk, Ad, name):
    return {None: spk.identity, "jacobi": spk.jacobi(Ad),
            "ilu": spk.ilu0(Ad)}[name]


def _minres(spk, O, A, b, tol, max_it, left=None, right=None, x0=None):
    Ad = spk.CsrMatrix.from_oracle(A)
    x = dv(spk, np.zeros(A.nrows) if x0 is None else x0)
    cfg = spk.MinresConfig(tol, max_it, True, spk.F64)
    s = spk.minres(dv(spk, b), x, cfg, Ad, _prec(spk, Ad, left),
                   _prec(spk, Ad, right))
    xr, sr = O.minres(A, b, x0, tol=tol, max_iterations=max_it, left=left,
                      right=right)
    return s, x.numpy(), sr, xr


def _qmr(spk, O, A, b, tol, max_it, left=None, right=None, x0=None):
    Ad = spk.CsrMatrix.from_oracle(A)
    At = Ad.transpose()
    x = dv(spk, np.zeros(A.nrows) if x0 is None else x0)
    cfg = spk.QmrConfig(tol, max_it, True, spk.F64)
    s = spk.qmr(dv(spk, b), x, cfg, Ad, At, _prec(spk, Ad, left),
                _prec(spk, Ad, right))
    xr, sr = O.qmr(A, b, x0, tol=tol, max_iterations=max_it, left=left,
                   right=right)
    return s, x.numpy(), sr, xr


def _indefinite4(O):
    return O.make_matrix(4, 4, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3],
                         [0, 1, 0, 1, 2, 1, 2, 3, 2, 3],
                         [2., -1., -1., -1., -1., -1., 3., -1., -1., -2.])


@pytest.mark.parametrize("case", ["tridiag", "grid16", "indefinite", "p2d"])
@pytest.mark.parametrize("left,right", [(None, None), ("jacobi", None),
                                        (None, "jacobi"), ("ilu", None)])
def test_minres_matches_oracle(O, spk, case, left, right):
    if case == "indefinite" and (left or right):
        pytest.skip("the diagonal or ILU(0) of an indefinite matrix is not "
                    "an SPD preconditioner, which MINRES needs")
    A = {"tridiag": O.tridiagonal_matrix(4, -1.0, 4.0, -1.0),
         "grid16": O.poisson_2d(4, 4, 5.0), "indefinite": _indefinite4(O),
         "p2d": O.poisson_2d(9, 8, 0.1)}[case]
    xt = np.arange(1, A.nrows + 1, dtype=float)
    b = O.spmv(A, xt)
    s, x, sr, xr = _minres(spk, O, A, b, 1e-10, 400, left, right)
    assert s.converged and sr.converged and not s.breakdown
    assert abs(s.computed_iterations - sr.computed_iterations) <= 1
    assert np.allclose(x, xt, rtol=0, atol=1e-6 * np.abs(xt).max())
    # histories agree until the estimates reach the rounding level of the
    # solve (the device's tree-ordered dots differ from the oracle's there)
    k = min(len(s.iteration_residuals), len(sr.iteration_residuals)) - 1
    assert np.allclose(s.iteration_residuals[:k], sr.iteration_residuals[:k],
                       rtol=1e-6, atol=1e-7 * np.linalg.norm(b))


@pytest.mark.parametrize("max_it", [1, 2, 3, 7, 12])
def test_minres_iterate_at_any_step(O, spk, max_it):
    """x after k steps (the d-vector update runs in the gated third pass,
    the vectors trade places by renaming) equals the oracle's."""
    A = O.poisson_2d(9, 8, 0.1)
    b = O.uniform_pm1(4, A.nrows)
    x0 = O.uniform_pm1(5, A.nrows)
    for left in (None, "jacobi", "ilu"):
        s, x, sr, xr = _minres(spk, O, A, b, 1e-30, max_it, left, None, x0)
        assert s.computed_iterations == max_it and not s.converged
        assert np.allclose(x, xr, rtol=0, atol=1e-11 * np.abs(xr).max())
        assert np.allclose(s.iteration_residuals, sr.iteration_residuals,
                           rtol=1e-9)


@pytest.mark.parametrize("case", ["tridiag", "nonsym", "grid16", "convdiff16",
                                  "convdiff2d"])
@pytest.mark.parametrize("left,right", [(None, None), ("jacobi", None),
                                        (None, "jacobi"), ("ilu", None),
                                        (None, "ilu")])
def test_qmr_matches_oracle(O, spk, case, left, right):
    if "ilu" in (left, right) and case not in ("tridiag", "grid16"):
        pytest.skip("the adjoint chain reuses M (qmr.hpp:300-303): an ILU(0) "
                    "of a nonsymmetric matrix makes the method inconsistent")
    A = {"tridiag": O.tridiagonal_matrix(4, -1.0, 4.0, -1.0),
         "nonsym": O.nonsymmetric_sample(), "grid16": O.poisson_2d(4, 4, 5.0),
         "convdiff16": O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0),
         "convdiff2d": O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)}[case]
    xt = np.arange(1, A.nrows + 1, dtype=float)
    b = O.spmv(A, xt)
    s, x, sr, xr = _qmr(spk, O, A, b, 1e-10, 600, left, right)
    assert s.converged and sr.converged and not s.breakdown
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2
    assert s.computed_iterations % 2 == 0
    assert O.true_relres(A, b, x) < 1e-8
    k = min(len(s.iteration_residuals), len(sr.iteration_residuals)) - 1
    assert np.allclose(s.iteration_residuals[:k], sr.iteration_residuals[:k],
                       rtol=1e-6, atol=1e-7 * np.linalg.norm(b))


@pytest.mark.parametrize("max_it", [2, 4, 6, 14])
def test_qmr_iterate_at_any_step(O, spk, max_it):
    A = O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(6, A.nrows)
    x0 = O.uniform_pm1(7, A.nrows)
    for left in (None, "jacobi"):
        s, x, sr, xr = _qmr(spk, O, A, b, 1e-30, max_it, left, None, x0)
        assert s.computed_iterations == max_it and not s.converged
        assert np.allclose(x, xr, rtol=0, atol=1e-10 * np.abs(xr).max())
        assert np.allclose(s.iteration_residuals, sr.iteration_residuals,
                           rtol=1e-8)


def test_zero_rhs_and_fp32(O, spk):
    # minres_test.cpp:198-217, qmr_test.cpp:295-328
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    Ad = spk.CsrMatrix.from_oracle(T)
    x = dv(spk, [5., 6., 7., 8.])
    s = spk.minres(dv(spk, np.zeros(4)), x, spk.MinresConfig(1e-12, 100), Ad)
    assert s.converged and np.all(x.numpy() == 0.0)
    x = dv(spk, [5., 6., 7., 8.])
    s = spk.qmr(dv(spk, np.zeros(4)), x, spk.QmrConfig(1e-12, 100), Ad,
                Ad.transpose())
    assert s.converged and np.all(x.numpy() == 0.0)
    # FP32 vectors, FP64 scalars
    A = O.convdiff_3d(10, 9, 8, 1.0, 0.3, 0.2, 0.1)
    b = O.uniform_pm1(3, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    x32 = dv(spk, np.zeros(A.nrows), np.float32)
    s = spk.qmr(dv(spk, b, np.float32), x32, spk.QmrConfig(1e-5, 600, False,
                                                           spk.F32),
                Ad, Ad.transpose(), spk.jacobi(Ad))
    assert s.converged
    assert O.true_relres(A, b, x32.numpy().astype(np.float64)) < 1e-4
    S = O.poisson_3d(10, 9, 8)
    Sd = spk.CsrMatrix.from_oracle(S)
    y32 = dv(spk, np.zeros(S.nrows), np.float32)
    s = spk.minres(dv(spk, b[:S.nrows], np.float32), y32,
                   spk.MinresConfig(1e-5, 600, False, spk.F32), Sd, spk.jacobi(Sd))
    assert s.converged
    assert O.true_relres(S, b[:S.nrows], y32.numpy().astype(np.float64)) < 1e-4


def test_medium_problems(O, spk):
    S = O.poisson_3d(18, 16, 14)
    b = O.uniform_pm1(8, S.nrows)
    s, x, sr, xr = _minres(spk, O, S, b, 1e-9, 2000, "jacobi")
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2
    assert O.true_relres(S, b, x) < 1e-8
    A = O.convdiff_3d(16, 14, 12, 1.0, 0.4, 0.2, 0.1)
    b = O.uniform_pm1(9, A.nrows)
    s, x, sr, xr = _qmr(spk, O, A, b, 1e-9, 2000, "jacobi")
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 4
    assert O.true_relres(A, b, x) < 1e-7
