"""Pins the CPU oracle's Krylov solvers: the reference's own solver tests
(conjugate_gradient_test.cpp, gmres_test.cpp, fgmres_test.cpp,
bicgstab_test.cpp, jacobi_test.cpp, ilu_test.cpp) plus invariants that a
wrong sign, index or dropped term would break: finite termination of CG and
GMRES, the restart rule of CG, the left-preconditioned residual GMRES
monitors, BiCGSTAB's matvec count, FGMRES == right-preconditioned GMRES
for a fixed preconditioner, and scipy reference solutions."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla


def grid16(O):
    return O.poisson_2d(4, 4, 5.0)


def problems(O):
    x4 = np.array([1., 2., 3., 4.])
    x16 = np.arange(1, 17, dtype=float)
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    cd = O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0)
    return {
        "identity": (O.make_matrix(4, 4, range(4), range(4), [1.] * 4),
                     x4.copy(), x4.copy(), 1e-12),
        "diagonal": (O.make_matrix(4, 4, range(4), range(4), [2., 3., 4., 5.]),
                     np.array([6., 12., 20., 30.]), np.array([3., 4., 5., 6.]),
                     1e-12),
        "tridiag": (tri, O.spmv(tri, x4), x4, 1e-12),
        "grid16": (grid16(O), O.spmv(grid16(O), x16), x16, 1e-10),
        "nonsym": (O.nonsymmetric_sample(),
                   O.spmv(O.nonsymmetric_sample(), x4), x4, 1e-12),
        "convdiff16": (cd, O.spmv(cd, x16), x16, 1e-10),
    }


def solve(O, method, A, b, tol, left=None, right=None, max_it=200,
          restart=20, x0=None):
    if method == "cg":
        return O.cg(A, b, x0, tol=tol, restart_iterations=restart,
                    max_iterations=max_it, left=left, right=right)
    if method == "gmres":
        return O.gmres(A, b, x0, tol=tol, restart_dimension=restart,
                       max_iterations=max_it, left=left, right=right)
    if method == "fgmres":
        return O.fgmres(A, b, x0, tol=tol, restart_dimension=restart,
                        max_iterations=max_it, preconditioner=left)
    return O.bicgstab(A, b, x0, tol=tol, max_iterations=max_it, left=left,
                      right=right)


@pytest.mark.parametrize("method,name", [
    (m, n) for m in ("cg",) for n in ("identity", "diagonal", "tridiag", "grid16")
] + [(m, n) for m in ("gmres", "fgmres", "bicgstab")
     for n in ("identity", "diagonal", "tridiag", "grid16", "nonsym",
               "convdiff16")])
def test_reference_solution_recovery(O, method, name):
    A, b, xt, tol = problems(O)[name]
    x, s = solve(O, method, A, b, tol)
    assert s.converged
    margin = 1e-10 if name in ("identity", "diagonal") else (
        1e-8 if A.nrows == 4 else 1e-6)
    assert np.allclose(x, xt, rtol=0, atol=margin)
    if name == "identity":
        assert s.computed_iterations <= 1


@pytest.mark.parametrize("method,left,right", [
    ("cg", "jacobi", None), ("cg", None, "jacobi"), ("gmres", "ilu", None),
    ("gmres", None, "ilu"), ("gmres", "milu", None), ("fgmres", "ilu", None),
    ("bicgstab", "ilu", None), ("bicgstab", None, "ilu")])
def test_preconditioned_tridiag_not_slower(O, method, left, right):
    # jacobi_test.cpp:97-143, ilu_test.cpp:109-156, *_test.cpp tridiag cases
    A, b, xt, tol = problems(O)["tridiag"]
    xp, sp = solve(O, method, A, b, tol, left, right)
    xu, su = solve(O, method, A, b, tol)
    assert sp.converged
    assert sp.computed_iterations <= su.computed_iterations
    assert np.allclose(xp, xt, atol=1e-8, rtol=0)


@pytest.mark.parametrize("method", ["cg", "gmres", "fgmres", "bicgstab"])
def test_zero_rhs_and_max_iterations(O, method):
    A = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    x, s = solve(O, method, A, np.zeros(4), 1e-12, x0=np.array([1., 2., 3., 4.]))
    assert s.converged and np.all(x == 0.0)
    b = O.spmv(grid16(O), np.arange(1, 17, dtype=float))
    mi = 2 if method == "bicgstab" else 1
    x, s = solve(O, method, grid16(O), b, 1e-14, max_it=mi, restart=10)
    assert not s.converged and s.computed_iterations == mi


def test_cg_terminates_in_n_steps(O):
    """Exact arithmetic: CG on an SPD n x n system terminates in <= n steps.
    In FP64 on a well-conditioned 16 x 16 system it reaches 1e-12."""
    A = grid16(O)
    b = O.uniform_pm1(3, 16)
    x, s = O.cg(A, b, tol=1e-12, restart_iterations=1000, max_iterations=16)
    assert s.converged and s.computed_iterations <= 16


def test_gmres_exact_at_step_n(O):
    """GMRES(m >= n) minimises over the whole space at step n."""
    A = O.convdiff_centered_2d(3, 3, 1.0, 0.7, 0.4, 0.1)
    b = O.uniform_pm1(4, 9)
    x, s = O.gmres(A, b, tol=1e-12, restart_dimension=9, max_iterations=9)
    assert s.converged and s.computed_iterations <= 9
    assert O.true_relres(A, b, x) < 1e-11


def test_cg_restart_resets_direction(O):
    """restart_iterations = 1 recomputes r and sets p = z every step
    (conjugate_gradient.hpp:208-230): that is preconditioned steepest
    descent, restated here independently."""
    A = O.poisson_2d(6, 6, 0.3)
    S = A.to_scipy()
    b = O.uniform_pm1(9, 36)
    d = 1.0 / S.diagonal()
    x_o, s_o = O.cg(A, b, tol=1e-30, restart_iterations=1, max_iterations=12,
                    left="jacobi")
    x = np.zeros(36)
    for _ in range(12):
        r = b - S @ x
        z = d * r
        alpha = (r @ z) / (z @ (S @ z))
        x = x + alpha * z
    assert np.allclose(x_o, x, rtol=1e-12, atol=1e-12)
    # and a restart period of 1 is much slower than none
    _, s1 = O.cg(A, b, tol=1e-10, restart_iterations=1, max_iterations=5000,
                 left="jacobi")
    _, s0 = O.cg(A, b, tol=1e-10, restart_iterations=5000,
                 max_iterations=5000, left="jacobi")
    assert s1.computed_iterations > 2 * s0.computed_iterations


def test_cg_residual_history_is_recursive_residual(O):
    A = O.poisson_2d(8, 8, 0.5)
    b = O.uniform_pm1(1, 64)
    x, s = O.cg(A, b, tol=1e-10, restart_iterations=1000, max_iterations=500,
                left="jacobi")
    assert len(s.iteration_residuals) == s.computed_iterations
    assert s.iteration_residuals[-1] == s.residual_norm
    assert s.residual_norm / np.linalg.norm(b) < 1e-10
    assert O.true_relres(A, b, x) < 1e-9


def test_gmres_monitors_left_preconditioned_residual(O):
    """With a left preconditioner GMRES tests |M^-1 (b - A x)| against
    |b| (gmres.hpp:202, 222-235), not the true residual."""
    A = O.convdiff_centered_2d(5, 5, 1.0, 0.3, 0.2, 5.0)
    # rescale rows so Jacobi changes the residual norm a lot
    S = A.to_scipy().tolil()
    for i in range(25):
        S[i, :] = S[i, :] * (1.0 + 9.0 * (i % 3))
    S = S.tocsr()
    B = O.Csr(25, 25, S.indptr.astype(np.int64), S.indices.astype(np.int64),
              S.data.copy())
    b = O.uniform_pm1(2, 25)
    x, s = O.gmres(B, b, tol=1e-6, restart_dimension=30, max_iterations=30,
                   left="jacobi")
    d = 1.0 / S.diagonal()
    prec_res = np.linalg.norm(d * (b - S @ x))
    true_res = np.linalg.norm(b - S @ x)
    assert s.residual_norm == pytest.approx(prec_res, rel=1e-6)
    assert abs(prec_res - true_res) > 0.5 * true_res


def test_gmres_history_monotone(O):
    # gmres_test.cpp:487-521, fgmres_test.cpp:381-415
    b = O.spmv(grid16(O), np.arange(1, 17, dtype=float))
    for f in (O.gmres, O.fgmres):
        x, s = f(grid16(O), b, tol=1e-10, restart_dimension=20,
                 max_iterations=200)
        h = s.iteration_residuals
        assert len(h) == s.computed_iterations
        assert np.all(h[1:] <= h[:-1] + 1e-14)


def test_fgmres_equals_right_preconditioned_gmres(O):
    """For a fixed preconditioner FGMRES and right-preconditioned GMRES
    build the same Krylov space (fgmres.hpp:250 vs gmres.hpp:250)."""
    A = O.convdiff_3d(6, 5, 4, 1.0, 0.4, 0.2, 0.1)
    b = O.uniform_pm1(8, A.nrows)
    xg, sg = O.gmres(A, b, tol=1e-10, restart_dimension=15,
                     max_iterations=300, right="ilu")
    xf, sf = O.fgmres(A, b, tol=1e-10, restart_dimension=15,
                      max_iterations=300, preconditioner="ilu")
    assert sg.computed_iterations == sf.computed_iterations
    assert np.allclose(sg.iteration_residuals, sf.iteration_residuals,
                       rtol=1e-8)
    assert np.allclose(xg, xf, rtol=1e-9, atol=1e-12)


def test_bicgstab_counts_matvecs(O):
    """computed_iterations counts matrix-vector products (bicgstab.hpp:249,
    289): a converged full step adds two, a half-step exit one."""
    A = O.make_matrix(4, 4, range(4), range(4), [1.] * 4)
    x, s = O.bicgstab(A, np.array([1., 2., 3., 4.]), tol=1e-12,
                      max_iterations=100)
    assert s.converged and s.computed_iterations == 1
    B = O.poisson_2d(7, 7, 0.2)
    b = O.uniform_pm1(3, 49)
    x, s = O.bicgstab(B, b, tol=1e-10, max_iterations=500)
    assert s.converged
    assert len(s.iteration_residuals) == s.computed_iterations


def test_bicgstab_breakdown_stops(O):
    A = O.tridiagonal_matrix(5, -1.0, 4.0, -1.0)
    xt = np.arange(1, 6, dtype=float)
    x, s = O.bicgstab(A, O.spmv(A, xt), xt, tol=1e-12, max_iterations=10)
    assert s.status == O.BREAKDOWN and not s.converged


@pytest.mark.parametrize("method,left", [("cg", "jacobi"), ("gmres", "ilu"),
                                          ("bicgstab", "ilu"),
                                          ("fgmres", "jacobi")])
def test_against_scipy_direct(O, method, left):
    if method == "cg":
        A = O.poisson_3d(8, 7, 6)
    else:
        A = O.convdiff_3d(8, 7, 6, 1.0, 0.5, 0.3, 0.1)
    b = O.uniform_pm1(11, A.nrows)
    x_ref = spla.spsolve(A.to_scipy().tocsc(), b)
    x, s = solve(O, method, A, b, 1e-10, left=left, restart=40, max_it=5000)
    assert s.converged
    assert np.linalg.norm(x - x_ref) <= 1e-7 * np.linalg.norm(x_ref)


def test_float_solvers(O):
    A = O.poisson_3d(8)
    b = O.uniform_pm1(12, A.nrows)
    x64, s64 = O.cg(A, b, tol=1e-5, restart_iterations=1000,
                    max_iterations=1000, left="jacobi")
    x32, s32 = O.cg(A, b, tol=1e-5, restart_iterations=1000,
                    max_iterations=1000, left="jacobi", dtype=np.float32)
    assert s32.converged and abs(s32.computed_iterations -
                                 s64.computed_iterations) <= 2
    assert np.allclose(x32, x64, rtol=0, atol=1e-4 * np.abs(x64).max())


@pytest.mark.parametrize("budget", [1, 3, 9])
def test_bicgstab_odd_budget_runs_a_whole_step(O, budget):
    """BiCGSTAB spends two matvecs per step and tests the budget before a
    step (bicgstab.hpp:176-178): an odd budget ends one matvec past it, with
    one residual-history entry per matvec (|s| and |r| of every step)."""
    A = O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(13, A.nrows)
    x, s = O.bicgstab(A, b, tol=1e-30, max_iterations=budget)
    assert s.computed_iterations == budget + 1
    assert len(s.iteration_residuals) == budget + 1
    assert np.all(np.isfinite(s.iteration_residuals))


def test_reference_call_pattern_gives_identical_results(O):
    """The timing variant of the operator slots (a fresh, zero-filled result
    vector per multiply and per triangular solve, copied to out, as
    sparse_blas.hpp:25 and ilu.hpp:173-179 do) computes the same numbers:
    iterates and histories are bit-identical."""
    A = O.convdiff_3d(9, 8, 7, 1.0, 0.3, 0.2, 0.1)
    b = O.uniform_pm1(5, A.nrows)
    for right in (None, "ilu"):
        x1, s1 = O.gmres(A, b, tol=1e-10, restart_dimension=20,
                         max_iterations=400, right=right)
        x2, s2 = O.gmres(A, b, tol=1e-10, restart_dimension=20,
                         max_iterations=400, right=right, ref_calls=True)
        assert np.array_equal(x1, x2)
        assert np.array_equal(s1.iteration_residuals, s2.iteration_residuals)
    P = O.poisson_2d(20, 20, 0.1)
    c = O.uniform_pm1(6, P.nrows)
    x1, s1 = O.cg(P, c, tol=1e-10, restart_iterations=500, max_iterations=500,
                  left="jacobi")
    x2, s2 = O.cg(P, c, tol=1e-10, restart_iterations=500, max_iterations=500,
                  left="jacobi", ref_calls=True)
    assert np.array_equal(x1, x2) and s1.computed_iterations == s2.computed_iterations
