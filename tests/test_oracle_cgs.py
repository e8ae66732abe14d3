"""The oracle's CGS (cgs.hpp:86-395) on the reference's own test cases
(cgs_test.cpp:54-254: identity, diagonal, tridiagonal, nonsymmetric sample,
16-node grid, 16-node convection-diffusion, zero rhs, iteration budget),
plus scipy solutions and the matvec count."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla


def cases(O):
    I4 = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [1., 1., 1., 1.])
    D4 = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    ns = O.nonsymmetric_sample()
    grid = O.poisson_2d(4, 4, 5.0)
    cd = O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0)
    x4, x16 = np.arange(1.0, 5.0), np.arange(1.0, 17.0)
    return [("identity", I4, np.array([1., 2., 3., 4.]), np.array([1., 2., 3., 4.]), 1e-12, 1e-10),
            ("diagonal", D4, np.array([6., 12., 20., 30.]), np.array([3., 4., 5., 6.]), 1e-12, 1e-10),
            ("tridiag", tri, O.spmv(tri, x4), x4, 1e-12, 1e-8),
            ("nonsym", ns, O.spmv(ns, x4), x4, 1e-12, 1e-8),
            ("grid16", grid, O.spmv(grid, x16), x16, 1e-10, 1e-6),
            ("convdiff16", cd, O.spmv(cd, x16), x16, 1e-10, 1e-6)]


@pytest.mark.parametrize("k", range(6))
def test_reference_cases(O, k):
    name, A, b, xt, tol, margin = cases(O)[k]
    x, s = O.cgs(A, b, tol=tol, max_iterations=200 if k >= 4 else 100)
    assert s.converged, name
    assert np.all(np.abs(x - xt) <= margin), name
    assert s.computed_iterations % 2 == 0          # two matvecs per step
    assert len(s.iteration_residuals) == s.computed_iterations // 2
    assert s.iteration_residuals[-1] / np.linalg.norm(b) < tol
    xs = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)


def test_zero_rhs_and_budget(O):
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    x, s = O.cgs(tri, np.zeros(4), np.ones(4), tol=1e-12, max_iterations=100)
    assert s.converged and np.all(x == 0.0)
    grid = O.poisson_2d(4, 4, 5.0)
    b = O.spmv(grid, np.arange(1.0, 17.0))
    x, s = O.cgs(grid, b, tol=1e-14, max_iterations=2)
    assert not s.converged and s.computed_iterations <= 2


@pytest.mark.parametrize("left,right", [("jacobi", None), (None, "jacobi"),
                                        ("ilu", None), (None, "ilu")])
def test_preconditioned(O, left, right):
    A = O.convdiff_centered_2d(8, 8, 1.0, 0.3, 0.3, 5.0)
    xt = np.arange(1.0, 65.0)
    b = O.spmv(A, xt)
    x0, s0 = O.cgs(A, b, tol=1e-10, max_iterations=400)
    x, s = O.cgs(A, b, tol=1e-10, max_iterations=400, left=left, right=right)
    assert s.converged and np.allclose(x, xt, atol=1e-6)
    assert s.computed_iterations <= s0.computed_iterations
