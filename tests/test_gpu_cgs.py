"""Device CGS against the oracle's CGS (pinned by the reference's own
cgs_test.cpp cases in test_oracle_cgs.py): same convergence, matvec counts
within one step (tree-ordered dot products), solutions to the reference
tests' margins, FP64 true residuals; FP32 vectors; both preconditioner
slots."""
import numpy as np
import pytest

from test_oracle_cgs import cases

pytestmark = pytest.mark.gpu


def dv(spk, a, dt=np.float64):
    return spk.DeviceVector.from_numpy(np.asarray(a, dtype=dt))


@pytest.mark.parametrize("k", range(6))
def test_reference_cases(O, spk, k):
    name, A, b, xt, tol, margin = cases(O)[k]
    mx = 200 if k >= 4 else 100
    Ad = spk.CsrMatrix.from_oracle(A)
    x = dv(spk, np.zeros(A.nrows))
    s = spk.cgs(dv(spk, b), x, spk.CgsConfig(tol, mx, True), Ad)
    xr, sr = O.cgs(A, b, tol=tol, max_iterations=mx)
    assert s.converged == sr.converged == True, name
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2, name
    assert np.all(np.abs(x.numpy() - xt) <= margin), name
    assert len(s.iteration_residuals) == s.computed_iterations // 2


def test_zero_rhs_budget_and_fp32(O, spk):
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    Ad = spk.CsrMatrix.from_oracle(tri)
    x = dv(spk, np.ones(4))
    s = spk.cgs(dv(spk, np.zeros(4)), x, spk.CgsConfig(1e-12, 100), Ad)
    assert s.converged and np.all(x.numpy() == 0.0)
    grid = O.poisson_2d(4, 4, 5.0)
    b = O.spmv(grid, np.arange(1.0, 17.0))
    Gd = spk.CsrMatrix.from_oracle(grid)
    x = dv(spk, np.zeros(16))
    s = spk.cgs(dv(spk, b), x, spk.CgsConfig(1e-14, 2), Gd)
    assert not s.converged and s.computed_iterations <= 2
    A = O.convdiff_centered_2d(10, 10, 1.0, 0.3, 0.3, 5.0)
    b = O.spmv(A, np.arange(1.0, 101.0))
    x = dv(spk, np.zeros(100), np.float32)
    s = spk.cgs(dv(spk, b, np.float32), x,
                spk.CgsConfig(1e-5, 400, False, spk.F32),
                spk.CsrMatrix.from_oracle(A))
    assert s.converged
    assert O.true_relres(A, b, x.numpy().astype(np.float64)) < 1e-4


@pytest.mark.parametrize("left,right", [("jacobi", None), (None, "jacobi"),
                                        ("ilu", None), (None, "ilu"),
                                        ("ilu", "jacobi")])
def test_preconditioned(O, spk, left, right):
    A = O.convdiff_centered_2d(12, 12, 1.0, 0.3, 0.3, 5.0)
    b = O.spmv(A, np.arange(1.0, 145.0))
    Ad = spk.CsrMatrix.from_oracle(A)
    mk = {"jacobi": spk.jacobi, "ilu": spk.ilu0, None: lambda _A: spk.identity}
    x = dv(spk, np.zeros(A.nrows))
    s = spk.cgs(dv(spk, b), x, spk.CgsConfig(1e-10, 600), Ad, mk[left](Ad),
                mk[right](Ad))
    xr, sr = O.cgs(A, b, tol=1e-10, max_iterations=600, left=left, right=right)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2
    assert O.true_relres(A, b, x.numpy()) < 1e-9


@pytest.mark.parametrize("budget", [2, 4, 6, 10])
@pytest.mark.parametrize("left,right", [(None, None), ("jacobi", None),
                                        (None, "ilu")])
def test_iterate_at_any_step(O, spk, budget, left, right):
    """CGS stopped by its matvec budget at any step returns the oracle's
    iterate and residual history (cgs.hpp:176-270): every pass that one
    step's scalars reach ran, and none of the next step's."""
    A = O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(15, A.nrows)
    x0 = O.uniform_pm1(16, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    mk = {"jacobi": spk.jacobi, "ilu": spk.ilu0, None: lambda _A: spk.identity}
    x = dv(spk, x0)
    s = spk.cgs(dv(spk, b), x, spk.CgsConfig(1e-30, budget, True), Ad,
                mk[left](Ad), mk[right](Ad))
    xr, sr = O.cgs(A, b, x0, tol=1e-30, max_iterations=budget, left=left,
                   right=right)
    assert not s.converged and s.computed_iterations == sr.computed_iterations
    assert np.allclose(x.numpy(), xr, rtol=0, atol=1e-10 * np.abs(xr).max())
    assert np.allclose(s.iteration_residuals, sr.iteration_residuals, rtol=1e-8)


def test_exact_start_is_a_breakdown(O, spk):
    """x0 exact: rho = <r^, r> = 0 and the first alpha's denominator is
    zero as well (the reference divides by it, cgs.hpp:224-226); the device
    stops with a breakdown before any update, like the oracle."""
    A = O.convdiff_centered_2d(6, 5, 1.0, 0.3, 0.2, 2.0)
    xt = O.uniform_pm1(3, A.nrows)
    b = O.spmv(A, xt)
    xr, sr = O.cgs(A, b, xt, tol=1e-12, max_iterations=50)
    assert sr.status != 0 and sr.computed_iterations == 0
    for left in (None, "jacobi"):
        Ad = spk.CsrMatrix.from_oracle(A)
        M = spk.jacobi(Ad) if left else spk.identity
        x = dv(spk, xt)
        s = spk.cgs(dv(spk, b), x, spk.CgsConfig(1e-12, 50), Ad, M)
        assert s.breakdown and not s.converged and s.computed_iterations == 0
        assert np.array_equal(x.numpy(), xt)
