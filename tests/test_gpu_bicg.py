"""Device transpose store and BiCG against the oracle (multiply_transpose and
bicg.hpp restated, pinned in test_oracle_bicg.py)."""
import numpy as np
import pytest

from test_oracle_cgs import cases

pytestmark = pytest.mark.gpu


def dv(spk, a, dt=np.float64):
    return spk.DeviceVector.from_numpy(np.asarray(a, dtype=dt))


@pytest.mark.parametrize("which", ["kat", "rect", "convdiff", "powerlaw"])
def test_transpose_spmv_bit_exact(O, spk, which):
    if which == "kat":
        A = O.make_matrix(3, 3, [0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [2., 3., 4., 5., 6.])
    elif which == "rect":
        A = O.make_matrix(2, 3, [0, 0, 1], [0, 2, 1], [1., 2., 3.])
    elif which == "convdiff":
        A = O.convdiff_3d(9, 8, 7, 1.0, 0.3, 0.2, 0.1)
    else:
        A = O.powerlaw_spd(4000, 5)
    At = spk.CsrMatrix.from_oracle(A).transpose()
    assert (At.nrows, At.ncols, At.nnz) == (A.ncols, A.nrows, A.nnz)
    x = O.uniform_pm1(3, A.nrows)
    y = spk.multiply(At, dv(spk, x)).numpy()
    yr = O.spmv_transpose(A, x)
    if which == "powerlaw":  # long columns are summed by several lanes
        absr = np.abs(A.dense().T) @ np.abs(x)
        assert np.all(np.abs(y - yr) <= 1e-12 * absr)
    else:
        assert np.array_equal(y.view(np.uint64), yr.view(np.uint64))


@pytest.mark.parametrize("k", range(6))
def test_reference_cases(O, spk, k):
    name, A, b, xt, tol, margin = cases(O)[k]
    mx = 200 if k >= 4 else 100
    Ad = spk.CsrMatrix.from_oracle(A)
    x = dv(spk, np.zeros(A.nrows))
    s = spk.bicg(dv(spk, b), x, spk.BicgConfig(tol, mx, True), Ad, Ad.transpose())
    xr, sr = O.bicg(A, b, tol=tol, max_iterations=mx)
    assert s.converged and sr.converged, name
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2, name
    assert np.all(np.abs(x.numpy() - xt) <= margin), name


@pytest.mark.parametrize("left,right", [("jacobi", None), (None, "jacobi"),
                                        ("ilu", "jacobi")])
def test_preconditioned_and_fp32(O, spk, left, right):
    A = O.convdiff_centered_2d(12, 12, 1.0, 0.3, 0.3, 5.0)
    b = O.spmv(A, np.arange(1.0, 145.0))
    Ad = spk.CsrMatrix.from_oracle(A)
    At = Ad.transpose()
    mk = {"jacobi": spk.jacobi, "ilu": spk.ilu0, None: lambda _A: spk.identity}
    x = dv(spk, np.zeros(A.nrows))
    s = spk.bicg(dv(spk, b), x, spk.BicgConfig(1e-10, 600), Ad, At, mk[left](Ad),
                 mk[right](Ad))
    xr, sr = O.bicg(A, b, tol=1e-10, max_iterations=600, left=left, right=right)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2
    assert O.true_relres(A, b, x.numpy()) < 1e-9
    x32 = dv(spk, np.zeros(A.nrows), np.float32)
    s = spk.bicg(dv(spk, b, np.float32), x32,
                 spk.BicgConfig(1e-5, 600, False, spk.F32), Ad, At, mk[left](Ad),
                 mk[right](Ad))
    assert s.converged
    assert O.true_relres(A, b, x32.numpy().astype(np.float64)) < 1e-4


@pytest.mark.parametrize("budget", [2, 4, 6, 10])
@pytest.mark.parametrize("left,right", [(None, None), ("jacobi", None),
                                        (None, "ilu")])
def test_iterate_at_any_step(O, spk, budget, left, right):
    """BiCG stopped by its matvec budget at any step returns the oracle's
    iterate and residual history (bicg.hpp:190-258)."""
    A = O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(17, A.nrows)
    x0 = O.uniform_pm1(18, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    mk = {"jacobi": spk.jacobi, "ilu": spk.ilu0, None: lambda _A: spk.identity}
    x = dv(spk, x0)
    s = spk.bicg(dv(spk, b), x, spk.BicgConfig(1e-30, budget, True), Ad,
                 Ad.transpose(), mk[left](Ad), mk[right](Ad))
    xr, sr = O.bicg(A, b, x0, tol=1e-30, max_iterations=budget, left=left,
                    right=right)
    assert not s.converged and s.computed_iterations == sr.computed_iterations
    assert np.allclose(x.numpy(), xr, rtol=0, atol=1e-10 * np.abs(xr).max())
    assert np.allclose(s.iteration_residuals, sr.iteration_residuals, rtol=1e-8)


def test_exact_start_is_a_breakdown(O, spk):
    """x0 exact: rho = <r~, z> = 0 and <p~, A p> = 0 (bicg.hpp:214-217);
    the device stops with a breakdown before any update, like the oracle."""
    A = O.convdiff_centered_2d(6, 5, 1.0, 0.3, 0.2, 2.0)
    xt = O.uniform_pm1(3, A.nrows)
    b = O.spmv(A, xt)
    xr, sr = O.bicg(A, b, xt, tol=1e-12, max_iterations=50)
    assert sr.status != 0 and sr.computed_iterations == 0
    for left in (None, "jacobi"):
        Ad = spk.CsrMatrix.from_oracle(A)
        M = spk.jacobi(Ad) if left else spk.identity
        x = dv(spk, xt)
        s = spk.bicg(dv(spk, b), x, spk.BicgConfig(1e-12, 50), Ad, Ad.transpose(), M)
        assert s.breakdown and not s.converged and s.computed_iterations == 0
        assert np.array_equal(x.numpy(), xt)
