"""The oracle's multiply_transpose (sparse_blas.hpp:36-52) and BiCG
(bicg.hpp:86-397) on the reference's own cases (sparse_blas_test.cpp
transpose KATs, bicg_test.cpp:56-370), dense A^T x, scipy solutions."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from test_oracle_cgs import cases


def test_transpose_kats(O):
    A = O.make_matrix(3, 3, [0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [2., 3., 4., 5., 6.])
    assert np.array_equal(O.spmv_transpose(A, [1., 2., 3.]), [17., 8., 21.])
    R = O.make_matrix(2, 3, [0, 0, 1], [0, 2, 1], [1., 2., 3.])
    assert np.array_equal(O.spmv_transpose(R, [1., 2.]), [1., 6., 2.])
    E = O.make_matrix(3, 4, [], [], [])
    assert np.array_equal(O.spmv_transpose(E, [1., 2., 3.]), np.zeros(4))
    P = O.powerlaw_spd(300, 2)
    x = O.uniform_pm1(1, 300)
    assert np.allclose(O.spmv_transpose(P, x), P.dense().T @ x, rtol=1e-13)


@pytest.mark.parametrize("k", range(6))
def test_reference_cases(O, k):
    name, A, b, xt, tol, margin = cases(O)[k]
    x, s = O.bicg(A, b, tol=tol, max_iterations=200 if k >= 4 else 100)
    assert s.converged, name
    assert np.all(np.abs(x - xt) <= margin), name
    assert s.computed_iterations % 2 == 0
    xs = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)


def test_zero_rhs_budget_preconditioned(O):
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    x, s = O.bicg(tri, np.zeros(4), np.ones(4), tol=1e-12, max_iterations=100)
    assert s.converged and np.all(x == 0.0)
    grid = O.poisson_2d(4, 4, 5.0)
    b = O.spmv(grid, np.arange(1.0, 17.0))
    _, s = O.bicg(grid, b, tol=1e-14, max_iterations=2)
    assert not s.converged and s.computed_iterations <= 2
    A = O.convdiff_centered_2d(8, 8, 1.0, 0.3, 0.3, 5.0)
    xt = np.arange(1.0, 65.0)
    b = O.spmv(A, xt)
    for left, right in [("jacobi", None), (None, "jacobi")]:
        x, s = O.bicg(A, b, tol=1e-10, max_iterations=400, left=left, right=right)
        assert s.converged and np.allclose(x, xt, atol=1e-6)
