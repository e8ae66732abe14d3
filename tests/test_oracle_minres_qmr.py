"""The oracle's MINRES (minres.hpp) and QMR (qmr.hpp) restatements, pinned by
the reference's own test cases (minres_test.cpp, qmr_test.cpp; their
incomplete-Cholesky preconditioner is replaced by ILU(0), which is the same
symmetric factorization L D L^T for these symmetric M-matrices) and by what
the mathematics fixes:

* MINRES step k minimises ||b - A x|| over x0 + K_k(A, r0): its iterate
  equals a dense least-squares solve over the Krylov basis, and its estimate
  |phi_bar| equals the true residual norm;
* for symmetric A with identity preconditioners the two-sided Lanczos of
  QMR collapses to the symmetric one, so QMR step k (2k matvecs) reproduces
  MINRES step k and |tau| equals |phi_bar|;
* QMR's quasi-residual bounds the true residual: ||r_k|| <= sqrt(k+1) |tau_k|.
"""
import numpy as np
import pytest


def _x16():
    return np.arange(1, 17, dtype=float)


def _indefinite4(O):
    # minres_test.cpp:48-62
    return O.make_matrix(4, 4, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3],
                         [0, 1, 0, 1, 2, 1, 2, 3, 2, 3],
                         [2., -1., -1., -1., -1., -1., 3., -1., -1., -2.])


# ------------------------------------------------------------- MINRES KATs

def test_minres_identity_diagonal_tridiag(O):
    # minres_test.cpp:70-144
    I4 = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [1.] * 4)
    b = np.array([1., 2., 3., 4.])
    x, s = O.minres(I4, b, tol=1e-12, max_iterations=100)
    assert s.converged and s.computed_iterations <= 1
    assert np.allclose(x, b, atol=1e-10)
    D = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    x, s = O.minres(D, np.array([6., 12., 20., 30.]), tol=1e-12,
                    max_iterations=100)
    assert s.converged and np.allclose(x, [3., 4., 5., 6.], atol=1e-10)
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    xt = np.array([1., 2., 3., 4.])
    x, s = O.minres(T, O.spmv(T, xt), tol=1e-12, max_iterations=100)
    assert s.converged and np.allclose(x, xt, atol=1e-8)


def test_minres_grid_indefinite_zero_rhs_max_iterations(O):
    # minres_test.cpp:146-241
    G = O.poisson_2d(4, 4, 5.0)
    x, s = O.minres(G, O.spmv(G, _x16()), tol=1e-10, max_iterations=200)
    assert s.converged and np.allclose(x, _x16(), atol=1e-6)
    A = _indefinite4(O)
    xt = np.array([1., 2., 3., 4.])
    x, s = O.minres(A, O.spmv(A, xt), tol=1e-10, max_iterations=100)
    assert s.converged and np.allclose(x, xt, atol=1e-6)
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    x, s = O.minres(T, np.zeros(4), np.array([5., 6., 7., 8.]), tol=1e-12,
                    max_iterations=100)
    assert s.converged and np.all(x == 0.0)
    x, s = O.minres(G, O.spmv(G, _x16()), tol=1e-14, max_iterations=1)
    assert not s.converged and s.computed_iterations == 1


def test_minres_monotone_history_and_preconditioned(O):
    # minres_test.cpp:243-456 (ILU(0) in place of incomplete Cholesky)
    G = O.poisson_2d(6, 5, 0.1)
    b = O.uniform_pm1(3, G.nrows)
    x, s = O.minres(G, b, tol=1e-12, max_iterations=200)
    h = np.array(s.iteration_residuals)
    assert s.converged and len(h) == s.computed_iterations
    assert np.all(h[1:] <= h[:-1] + 1e-14)
    for left, right in (("ilu", None), (None, "ilu"), ("jacobi", None)):
        xp, sp = O.minres(G, b, tol=1e-12, max_iterations=200, left=left,
                          right=right)
        assert sp.converged
        if left != "jacobi":  # (the estimate is in the M^-1 norm)
            assert sp.computed_iterations <= s.computed_iterations
        assert O.true_relres(G, b, xp) < 1e-9


@pytest.mark.parametrize("k", [1, 2, 4, 7])
def test_minres_is_krylov_least_squares(O, k):
    A = _indefinite4(O) if k <= 2 else O.poisson_2d(6, 5, 0.1)
    n = A.nrows
    b = O.uniform_pm1(5, n)
    x, s = O.minres(A, b, tol=1e-30, max_iterations=k)
    Ad = A.dense()
    K = np.column_stack([np.linalg.matrix_power(Ad, j) @ b for j in range(k)])
    c, *_ = np.linalg.lstsq(Ad @ K, b, rcond=None)
    xl = K @ c
    assert np.allclose(x, xl, rtol=0, atol=1e-9 * np.abs(xl).max())
    true = np.linalg.norm(b - Ad @ x)
    assert abs(s.residual_norm - true) <= 1e-9 * np.linalg.norm(b)


# ---------------------------------------------------------------- QMR KATs

def test_qmr_reference_cases(O):
    # qmr_test.cpp:56-290
    I4 = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [1.] * 4)
    b = np.array([1., 2., 3., 4.])
    x, s = O.qmr(I4, b, tol=1e-12, max_iterations=100)
    assert s.converged and np.allclose(x, b, atol=1e-10)
    D = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    x, s = O.qmr(D, np.array([6., 12., 20., 30.]), tol=1e-12, max_iterations=100)
    assert s.converged and np.allclose(x, [3., 4., 5., 6.], atol=1e-10)
    xt = np.array([1., 2., 3., 4.])
    for A in (O.tridiagonal_matrix(4, -1.0, 4.0, -1.0), O.nonsymmetric_sample()):
        x, s = O.qmr(A, O.spmv(A, xt), tol=1e-12, max_iterations=100)
        assert s.converged and np.allclose(x, xt, atol=1e-8)
    for A in (O.poisson_2d(4, 4, 5.0),
              O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0)):
        x, s = O.qmr(A, O.spmv(A, _x16()), tol=1e-10, max_iterations=200)
        assert s.converged and np.allclose(x, _x16(), atol=1e-6)
        assert s.computed_iterations % 2 == 0


def test_qmr_zero_rhs_max_iterations_preconditioned(O):
    # qmr_test.cpp:295-556 (ILU(0) in place of incomplete Cholesky)
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    x, s = O.qmr(T, np.zeros(4), np.array([5., 6., 7., 8.]), tol=1e-12,
                 max_iterations=100)
    assert s.converged and np.all(x == 0.0)
    G = O.poisson_2d(4, 4, 5.0)
    x, s = O.qmr(G, O.spmv(G, _x16()), tol=1e-14, max_iterations=2)
    assert not s.converged and s.computed_iterations <= 2
    for left, right in (("ilu", None), (None, "ilu")):
        x, s = O.qmr(G, O.spmv(G, _x16()), tol=1e-8, max_iterations=400,
                     left=left, right=right)
        assert s.converged and np.allclose(x, _x16(), atol=1e-4)


@pytest.mark.parametrize("k", [1, 2, 3, 6])
def test_qmr_equals_minres_on_symmetric(O, k):
    A = O.poisson_2d(6, 5, 0.1)
    b = O.uniform_pm1(8, A.nrows)
    xm, sm = O.minres(A, b, tol=1e-30, max_iterations=k)
    xq, sq = O.qmr(A, b, tol=1e-30, max_iterations=2 * k)
    assert sq.computed_iterations == 2 * k
    assert np.allclose(xq, xm, rtol=0, atol=1e-10 * np.abs(xm).max())
    assert abs(sq.residual_norm - sm.residual_norm) <= 1e-10 * np.linalg.norm(b)


def test_qmr_quasi_residual_bound(O):
    A = O.convdiff_centered_2d(7, 6, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(9, A.nrows)
    Ad = A.dense()
    for k in range(1, 12):
        x, s = O.qmr(A, b, tol=1e-30, max_iterations=2 * k)
        r = np.linalg.norm(b - Ad @ x)
        assert r <= np.sqrt(k + 1) * s.residual_norm * (1 + 1e-9)
    x, s = O.qmr(A, b, tol=1e-10, max_iterations=400)
    assert s.converged and O.true_relres(A, b, x) < 1e-8
