"""The oracle's approximate ILU(0) apply (k Jacobi sweeps per triangle; no
reference counterpart, it replaces the exact triangular solves of
ilu.hpp:173-179) pinned by what the mathematics fixes:

* the strict triangles are nilpotent, so k >= the dependency depth gives
  the exact ilu_apply (the oracle's forward_solve/backward_solve,
  themselves pinned against ilu_test.cpp);
* k = 0 is the diagonal scaling r / diag(U);
* for any k the result is the truncated Neumann polynomial of each factor,
  evaluated here with dense matrix powers."""
import numpy as np
import pytest


def _mats(O):
    return {
        "poisson2d": O.poisson_2d(6, 5, 0.1),
        "convdiff2d": O.convdiff_centered_2d(5, 4, 1.0, 0.3, 0.2, 2.0),
        "convdiff3d": O.convdiff_3d(4, 3, 3, 1.0, 0.4, 0.2, 0.1),
        "tridiag": O.tridiagonal_matrix(9, -1.0, 4.0, -1.5),
        "nonsym": O.nonsymmetric_sample(),
    }


def _depth(F, lower):
    """Longest dependency chain of a triangular factor (edges)."""
    n = F.nrows
    lev = np.zeros(n, np.int64)
    order = range(n) if lower else range(n - 1, -1, -1)
    for i in order:
        for p in range(F.row_ptr[i], F.row_ptr[i + 1]):
            j = F.col_ind[p]
            if j != i:
                lev[i] = max(lev[i], lev[j] + 1)
    return int(lev.max())


def _exact(O, L, U, r):
    return O.backward_solve(U, O.forward_solve(L, r))


def _neumann(L, U, k, r):
    Ld, Ud = L.dense(), U.dense()
    n = len(r)
    Ls = np.tril(Ld, -1)
    d = np.diag(Ud)
    N = np.triu(Ud, 1) / d[:, None]
    # y = sum_{j<=k} (-L_s)^j r;  z = sum_{j<=k} (-D^-1 U_s)^j D^-1 y
    y, P = np.zeros(n), np.eye(n)
    for _ in range(k + 1):
        y += P @ r
        P = P @ (-Ls)
    z, P = np.zeros(n), np.eye(n)
    for _ in range(k + 1):
        z += P @ (y / d)
        P = P @ (-N)
    return z


@pytest.mark.parametrize("name", ["poisson2d", "convdiff2d", "convdiff3d",
                                  "tridiag", "nonsym"])
def test_exact_at_dependency_depth(O, name):
    A = _mats(O)[name]
    L, U = O.ilu0(A)
    r = O.uniform_pm1(5, A.nrows)
    ex = _exact(O, L, U, r)
    k = max(_depth(L, True), _depth(U, False))
    for kk in (k, k + 3):
        z = O.ilu_jacobi_apply(L, U, kk, r)
        assert np.allclose(z, ex, rtol=0, atol=1e-13 * np.abs(ex).max())
    if k > 1:  # and not before: the truncation is visible
        z = O.ilu_jacobi_apply(L, U, k - 1, r)
        assert np.abs(z - ex).max() > 1e-12 * np.abs(ex).max()


def test_zero_sweeps_is_diagonal_scaling(O):
    A = O.poisson_2d(6, 5, 0.1)
    L, U = O.ilu0(A)
    r = O.uniform_pm1(6, A.nrows)
    d = np.array([U.vals[U.row_ptr[i]] for i in range(A.nrows)])
    assert np.array_equal(O.ilu_jacobi_apply(L, U, 0, r), r / d)


@pytest.mark.parametrize("k", [1, 2, 3, 5])
@pytest.mark.parametrize("name", ["poisson2d", "convdiff3d", "nonsym"])
def test_truncated_neumann_polynomial(O, name, k):
    A = _mats(O)[name]
    L, U = O.ilu0(A)
    r = O.uniform_pm1(7, A.nrows)
    z = O.ilu_jacobi_apply(L, U, k, r)
    ref = _neumann(L, U, k, r)
    assert np.allclose(z, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_float32_follows_float64(O):
    A = O.convdiff_3d(5, 4, 4, 1.0, 0.4, 0.2, 0.1)
    L, U = O.ilu0(A)
    r = O.uniform_pm1(8, A.nrows)
    z64 = O.ilu_jacobi_apply(L, U, 3, r)
    z32 = O.ilu_jacobi_apply(L, U, 3, r.astype(np.float32))
    assert z32.dtype == np.float32
    assert np.allclose(z32, z64, rtol=0, atol=1e-5 * np.abs(z64).max())


def test_gmres_with_approximate_ilu(O):
    """A fixed linear preconditioner: GMRES converges, between the exact
    ILU(0) and Jacobi in iterations."""
    A = O.convdiff_3d(10, 10, 10, 1.0, 0.4, 0.2, 0.1)
    b = O.uniform_pm1(9, A.nrows)
    its = {}
    for name, left in (("ilu", "ilu"), ("jac3", ("ilu_jac", 3)),
                       ("jacobi", "jacobi")):
        x, s = O.gmres(A, b, tol=1e-10, restart_dimension=50,
                       max_iterations=500, left=left)
        assert s.converged and O.true_relres(A, b, x) < 1e-9
        its[name] = s.computed_iterations
    assert its["ilu"] <= its["jac3"] < its["jacobi"], its
