"""GPU parity of the device Krylov solvers against the oracle and the
reference's own solver tests (conjugate_gradient_test.cpp, gmres_test.cpp,
fgmres_test.cpp, bicgstab_test.cpp, jacobi_test.cpp, ilu_test.cpp).

Iteration counts: the device reduces dot products in a different (tree)
order than std::inner_product, so counts are compared within +-1 of the
oracle's, and every converged solve is checked on its FP64 true residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dv(spk, a, dtype=np.float64):
    return spk.DeviceVector.from_numpy(np.asarray(a, dtype=dtype))


def grid16(O):
    return O.poisson_2d(4, 4, 5.0)  # make_grid_16: matgen_test.cpp:348-379


def _problems(O):
    x4 = np.array([1., 2., 3., 4.])
    x16 = np.arange(1, 17, dtype=float)
    diag = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    ident = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [1., 1., 1., 1.])
    tri = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    return {
        "identity": (ident, np.array([1., 2., 3., 4.]), 1e-12),
        "diagonal": (diag, np.array([6., 12., 20., 30.]), 1e-12),
        "tridiag": (tri, O.spmv(tri, x4), 1e-12),
        "grid16": (grid16(O), O.spmv(grid16(O), x16), 1e-10),
        "nonsym": (O.nonsymmetric_sample(),
                   O.spmv(O.nonsymmetric_sample(), x4), 1e-12),
        "convdiff16": (O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0),
                       O.spmv(O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0),
                              x16), 1e-10),
    }


SYM = ["identity", "diagonal", "tridiag", "grid16"]
ALL = SYM + ["nonsym", "convdiff16"]


def _run(spk, O, method, A, b, tol, left=None, right=None, max_it=200,
         restart=50, collect=True, x0=None, precision=None):
    Ad = spk.CsrMatrix.from_oracle(A)
    prec = {"jacobi": spk.jacobi, "ilu": spk.ilu0, "milu": spk.milu0,
            None: lambda _A: spk.identity}
    ML, MR = prec[left](Ad), prec[right](Ad)
    p = spk.F64 if precision is None else precision
    dt = np.float32 if p == spk.F32 else np.float64
    bd = dv(spk, b, dt)
    xd = dv(spk, np.zeros(A.nrows) if x0 is None else x0, dt)
    if method == "cg":
        cfg = spk.CGConfig(tol, restart, max_it, collect, p)
        s = spk.conjugate_gradient(bd, xd, cfg, Ad, ML, MR)
        xr, sr = O.cg(A, b, x0, tol=tol, restart_iterations=restart,
                      max_iterations=max_it, left=left, right=right)
    elif method == "gmres":
        cfg = spk.GmresConfig(tol, restart, max_it, collect, p)
        s = spk.gmres(bd, xd, cfg, Ad, ML, MR)
        xr, sr = O.gmres(A, b, x0, tol=tol, restart_dimension=restart,
                         max_iterations=max_it, left=left, right=right)
    elif method == "fgmres":
        cfg = spk.GmresConfig(tol, restart, max_it, collect, p)
        s = spk.fgmres(bd, xd, cfg, Ad, ML)
        xr, sr = O.fgmres(A, b, x0, tol=tol, restart_dimension=restart,
                          max_iterations=max_it, preconditioner=left)
    else:
        cfg = spk.BicgstabConfig(tol, max_it, collect, p)
        s = spk.bicgstab(bd, xd, cfg, Ad, ML, MR)
        xr, sr = O.bicgstab(A, b, x0, tol=tol, max_iterations=max_it,
                            left=left, right=right)
    return s, xd.numpy(), sr, xr


@pytest.mark.parametrize("method,name", [("cg", n) for n in SYM] +
                         [(m, n) for m in ("gmres", "fgmres", "bicgstab")
                          for n in ALL])
def test_unpreconditioned(O, spk, method, name):
    A, b, tol = _problems(O)[name]
    s, x, sr, xr = _run(spk, O, method, A, b, tol, restart=20)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 1
    assert np.allclose(x, xr, rtol=0, atol=1e-8)
    if name == "identity":
        assert s.computed_iterations <= 1


@pytest.mark.parametrize("method,left,right", [
    ("cg", "jacobi", None), ("cg", None, "jacobi"),
    ("gmres", "jacobi", None), ("gmres", None, "jacobi"),
    ("gmres", "ilu", None), ("gmres", None, "ilu"), ("gmres", "milu", None),
    ("fgmres", "jacobi", None), ("fgmres", "ilu", None),
    ("bicgstab", "jacobi", None), ("bicgstab", None, "ilu"),
    ("bicgstab", "ilu", "jacobi"),
])
@pytest.mark.parametrize("name", ["tridiag", "grid16", "convdiff16"])
def test_preconditioned(O, spk, method, left, right, name):
    if method == "cg" and name == "convdiff16":
        pytest.skip("CG needs a symmetric matrix")
    A, b, tol = _problems(O)[name]
    s, x, sr, xr = _run(spk, O, method, A, b, tol, left, right, restart=20)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 1
    assert np.allclose(x, xr, rtol=0, atol=1e-7)
    if "ilu" in (left, right) and name == "tridiag":
        # preconditioned <= unpreconditioned (ilu_test.cpp:148-152)
        su, *_ = _run(spk, O, method, A, b, tol, restart=20)
        assert s.computed_iterations <= su.computed_iterations


@pytest.mark.parametrize("method", ["cg", "gmres", "fgmres", "bicgstab"])
def test_zero_rhs(O, spk, method):
    A = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    s, x, *_ = _run(spk, O, method, A, np.zeros(4), 1e-12,
                    x0=np.array([5., 6., 7., 8.]))
    assert s.converged and np.all(x == 0.0)


@pytest.mark.parametrize("method,max_it,expect", [
    ("cg", 1, 1), ("gmres", 1, 1), ("fgmres", 1, 1), ("bicgstab", 2, 2)])
def test_max_iterations(O, spk, method, max_it, expect):
    # conjugate_gradient_test.cpp:377-400, gmres_test.cpp:462-485,
    # bicgstab_test.cpp:392-414
    A = grid16(O)
    b = O.spmv(A, np.arange(1, 17, dtype=float))
    s, *_ = _run(spk, O, method, A, b, 1e-14, max_it=max_it, restart=10)
    assert not s.converged
    assert s.computed_iterations == expect


@pytest.mark.parametrize("method", ["gmres", "fgmres"])
def test_residual_history_monotone(O, spk, method):
    # gmres_test.cpp:487-521, fgmres_test.cpp:381-415
    A = grid16(O)
    b = O.spmv(A, np.arange(1, 17, dtype=float))
    s, x, sr, xr = _run(spk, O, method, A, b, 1e-10, restart=20)
    h = np.array(s.iteration_residuals)
    assert len(h) == s.computed_iterations
    assert np.all(h[1:] <= h[:-1] + 1e-14)
    assert np.allclose(h, sr.iteration_residuals, rtol=1e-9, atol=1e-13)


def test_cg_restart_semantics(O, spk):
    """restart_iterations resets the search direction
    (conjugate_gradient.hpp:208-230): with restart 5 the iteration count and
    residual history follow the oracle's."""
    A = O.poisson_2d(24, 24, 0.05)
    b = O.uniform_pm1(3, A.nrows)
    s, x, sr, xr = _run(spk, O, "cg", A, b, 1e-9, left="jacobi", restart=5,
                        max_it=5000)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 2
    k = min(50, len(s.iteration_residuals), len(sr.iteration_residuals))
    assert np.allclose(s.iteration_residuals[:k], sr.iteration_residuals[:k],
                       rtol=1e-6)


def test_bicgstab_breakdown_reported(O, spk):
    """x0 exact: rho_new = <r^, r> = 0.  The reference returns from the step
    without consuming budget and loops forever (bicgstab.hpp:203-204); the
    device solver stops and reports the breakdown."""
    A = O.tridiagonal_matrix(6, -1.0, 4.0, -1.0)
    xt = np.arange(1, 7, dtype=float)
    s, x, *_ = _run(spk, O, "bicgstab", A, O.spmv(A, xt), 1e-12, x0=xt)
    assert s.breakdown and not s.converged
    assert np.array_equal(x, xt)


@pytest.mark.parametrize("copies", [1, 1_100_000])  # hand-off / finisher path
def test_bicgstab_omega_breakdown_keeps_half_step(spk, copies):
    """omega = <t,s>/<t,t> = 0 exactly: copies of the 2x2 block
    [[1, 1], [-1, 0]] with b = (1, 0) give alpha = 1, s = (0, 1) and
    t = A s = (1, 0), so <t, s> = 0 in the first step.  The reference
    returns before updating x and then divides by omega (bicgstab.hpp:222,
    :206); the device solver keeps the valid half step (x += alpha p^,
    r = s), stops with a breakdown, and reports |s|, which the returned x
    satisfies.  1.1M copies (2.2M rows, > 4096 SpMV row blocks) take the
    finisher path, one copy the one-kernel hand-off."""
    n = 2 * copies
    rp = np.zeros(n + 1, np.int64)
    rp[1::2] = np.arange(copies) * 3 + 2
    rp[2::2] = np.arange(1, copies + 1) * 3
    base = np.arange(copies) * 2
    ci = np.stack([base, base + 1, base], axis=1).ravel()
    vals = np.tile([1.0, 1.0, -1.0], copies)
    A = spk.CsrMatrix.from_host(n, n, rp, ci, vals)
    b = np.tile([1.0, 0.0], copies)
    x = spk.DeviceVector.zeros(n)
    s = spk.bicgstab(spk.DeviceVector.from_numpy(b), x,
                     spk.BicgstabConfig(1e-10, 100), A)
    assert s.breakdown and not s.converged
    assert s.computed_iterations == 2
    xh = x.numpy()
    assert np.array_equal(xh, np.tile([1.0, 0.0], copies))
    x0, x1 = xh[0::2], xh[1::2]
    r = np.stack([1.0 - (x0 + x1), x0], axis=1).ravel()  # b - A x per block
    assert np.array_equal(r, np.tile([0.0, 1.0], copies))
    assert s.residual_norm == np.sqrt(float(copies))  # |s| = |b - A x|


def test_gmres_exact_start(O, spk):
    """beta = 0 at a restart (x0 exact) is reported as converged instead of
    the reference's division by zero (gmres.hpp:229-231)."""
    A = O.nonsymmetric_sample()
    xt = np.array([1., 2., 3., 4.])
    s, x, *_ = _run(spk, O, "gmres", A, O.spmv(A, xt), 1e-12, x0=xt)
    assert s.converged and np.array_equal(x, xt)


@pytest.mark.parametrize("method,left", [("cg", "jacobi"), ("gmres", "ilu"),
                                          ("bicgstab", "ilu")])
def test_medium_problems(O, spk, method, left):
    if method == "cg":
        A = O.poisson_3d(24, 20, 16)
    else:
        A = O.convdiff_3d(20, 16, 12, 1.0, 0.4, 0.2, 0.1)
    b = O.uniform_pm1(42, A.nrows)
    s, x, sr, xr = _run(spk, O, method, A, b, 1e-8, left=left, restart=30,
                        max_it=3000)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= max(
        2, sr.computed_iterations // 50)
    # gmres with a left preconditioner tests the preconditioned residual
    # (gmres.hpp:202); the true residual is then checked against the
    # oracle's own solution
    rel = O.true_relres(A, b, x)
    rel_ref = O.true_relres(A, b, xr)
    assert rel <= max(1e-8, 10 * rel_ref)


@pytest.mark.parametrize("method", ["cg", "gmres", "bicgstab"])
def test_fp32_vectors(O, spk, method):
    """precision = SPK_F32: FP32 vectors/SpMV, FP64 dots and scalars."""
    if method == "cg":
        A, left = O.poisson_3d(16, 16, 16), "jacobi"
    else:
        A, left = O.convdiff_3d(16, 12, 10, 1.0, 0.4, 0.2, 0.1), "ilu"
    b = O.uniform_pm1(5, A.nrows)
    s, x, sr, xr = _run(spk, O, method, A, b, 1e-5, left=left, restart=30,
                        max_it=2000, precision=spk.F32)
    assert s.converged
    assert O.true_relres(A, b, x.astype(np.float64)) < 1e-4


def test_powerlaw_jacobi_pcg(O, spk):
    A = O.powerlaw_spd(20000, 11)
    b = O.uniform_pm1(42, A.nrows)
    s, x, sr, xr = _run(spk, O, "cg", A, b, 1e-8, left="jacobi",
                        restart=100000, max_it=2000)
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 1
    assert O.true_relres(A, b, x) < 1e-8


@pytest.mark.parametrize("method", ["gmres", "fgmres"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cgs2_orthogonalization(O, spk, method, prec):
    """Classical Gram-Schmidt twice gives the MGS Krylov basis up to
    rounding: same convergence, iteration counts within a step or two."""
    A = O.convdiff_3d(14, 12, 10, 1.0, 0.3, 0.2, 0.1)
    b = O.uniform_pm1(11, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    F = spk.ilu0(Ad)
    p = spk.F32 if prec == "f32" else spk.F64
    dt = np.float32 if prec == "f32" else np.float64
    tol = 1e-5 if prec == "f32" else 1e-10
    its = []
    for ortho in (spk.ORTHO_MGS, spk.ORTHO_CGS2):
        x = dv(spk, np.zeros(A.nrows), dt)
        cfg = spk.GmresConfig(tol, 20, 600, True, p, ortho)
        if method == "gmres":
            s = spk.gmres(dv(spk, b, dt), x, cfg, Ad, spk.identity, F)
        else:
            s = spk.fgmres(dv(spk, b, dt), x, cfg, Ad, F)
        assert s.converged
        rel = O.true_relres(A, b, x.numpy().astype(np.float64))
        assert rel < (1e-4 if prec == "f32" else 1e-9)
        its.append(s.computed_iterations)
    assert abs(its[0] - its[1]) <= 2, its
    if prec == "f64":
        if method == "gmres":
            _, sr = O.gmres(A, b, tol=tol, restart_dimension=20,
                            max_iterations=600, right="ilu")
        else:
            _, sr = O.fgmres(A, b, tol=tol, restart_dimension=20,
                             max_iterations=600, preconditioner="ilu")
        assert abs(its[1] - sr.computed_iterations) <= 2


@pytest.mark.parametrize("method", ["gmres", "fgmres"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_cgs1_orthogonalization(O, spk, method, prec):
    """Classical Gram-Schmidt once (SPK_ORTHO_CGS1, no reference
    counterpart): on this well-conditioned preconditioned problem the basis
    stays orthogonal enough that the counts match MGS within a step or two,
    and the true residual agrees with the recursive estimate."""
    A = O.convdiff_3d(14, 12, 10, 1.0, 0.3, 0.2, 0.1)
    b = O.uniform_pm1(11, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    F = spk.ilu0(Ad)
    p = spk.F32 if prec == "f32" else spk.F64
    dt = np.float32 if prec == "f32" else np.float64
    tol = 1e-5 if prec == "f32" else 1e-10
    its = []
    for ortho in (spk.ORTHO_MGS, spk.ORTHO_CGS1):
        x = dv(spk, np.zeros(A.nrows), dt)
        cfg = spk.GmresConfig(tol, 20, 600, True, p, ortho)
        if method == "gmres":
            s = spk.gmres(dv(spk, b, dt), x, cfg, Ad, spk.identity, F)
        else:
            s = spk.fgmres(dv(spk, b, dt), x, cfg, Ad, F)
        assert s.converged
        rel = O.true_relres(A, b, x.numpy().astype(np.float64))
        assert rel < (1e-4 if prec == "f32" else 1e-9)
        its.append(s.computed_iterations)
    assert abs(its[0] - its[1]) <= 2, its
    with pytest.raises(spk.InvalidArgument, match="orthogonalization"):
        spk.gmres(dv(spk, b, dt), dv(spk, np.zeros(A.nrows), dt),
                  spk.GmresConfig(tol, 20, 600, True, p, 3), Ad, spk.identity, F)


def test_ir_with_cgs1(O, spk):
    A = O.convdiff_3d(20, 16, 12, 1.0, 0.25, 0.12, 0.06)
    b = O.uniform_pm1(4, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.IrConfig(1e-10, 20, 1e-4, 40, 400, spk.IR_GMRES, spk.ORTHO_CGS1)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, cfg, Ad, spk.ilu0(Ad))
    assert s.converged and O.true_relres(A, b, x.numpy()) < 1e-10


def test_ir_with_cgs2(O, spk):
    A = O.convdiff_3d(20, 16, 12, 1.0, 0.25, 0.12, 0.06)
    b = O.uniform_pm1(4, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.IrConfig(1e-10, 20, 1e-4, 40, 400, spk.IR_GMRES, spk.ORTHO_CGS2)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, cfg, Ad, spk.ilu0(Ad))
    assert s.converged and O.true_relres(A, b, x.numpy()) < 1e-10


@pytest.mark.parametrize("max_it", [1, 4, 5, 6, 11, 23])
@pytest.mark.parametrize("panels", [1, 3])
def test_cg_jacobi_iterate_after_max_iterations(O, spk, max_it, panels):
    """Jacobi-PCG applies the x update in the direction-update pass, which
    is gated on the step count (it must also run in the step that stops
    the loop): a solve stopped at any step, with restarts every 5 steps,
    returns the oracle's iterate (conjugate_gradient.hpp:208-292)."""
    A = O.poisson_2d(20, 17, 0.05)
    b = O.uniform_pm1(11, A.nrows)
    x0 = O.uniform_pm1(12, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    Ad.set_panels(panels)
    bd, xd = dv(spk, b), dv(spk, x0)
    cfg = spk.CGConfig(1e-30, 5, max_it, True, spk.F64)
    s = spk.conjugate_gradient(bd, xd, cfg, Ad, spk.jacobi(Ad), spk.identity)
    xr, sr = O.cg(A, b, x0, tol=1e-30, restart_iterations=5,
                  max_iterations=max_it, left="jacobi", right=None)
    assert not s.converged and s.computed_iterations == max_it
    x = xd.numpy()
    assert np.allclose(x, xr, rtol=0, atol=1e-11 * np.abs(xr).max())
    assert np.allclose(s.iteration_residuals, sr.iteration_residuals,
                       rtol=1e-9)
    # FP32 vectors: same recurrence, FP32 roundings
    xf = dv(spk, x0, np.float32)
    cfg32 = spk.CGConfig(1e-30, 5, max_it, False, spk.F32)
    s32 = spk.conjugate_gradient(dv(spk, b, np.float32), xf, cfg32, Ad,
                                 spk.jacobi(Ad), spk.identity)
    assert s32.computed_iterations == max_it
    assert np.allclose(xf.numpy(), xr, rtol=0, atol=1e-4 * np.abs(xr).max())


@pytest.mark.parametrize("k", [2, 3])
def test_gmres_with_approximate_ilu(O, spk, k):
    """GMRES(50) with the Jacobi-sweep ILU(0) apply (a fixed linear
    preconditioner) follows the oracle's GMRES with the same operator."""
    A = O.convdiff_3d(14, 12, 10, 1.0, 0.4, 0.2, 0.1)
    b = O.uniform_pm1(9, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    F = spk.ilu0(Ad)
    F.jacobi_sweeps = k
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.GmresConfig(1e-10, 50, 500, True, spk.F64)
    s = spk.gmres(spk.DeviceVector.from_numpy(b), x, cfg, Ad, F, spk.identity)
    xr, sr = O.gmres(A, b, tol=1e-10, restart_dimension=50, max_iterations=500,
                     left=("ilu_jac", k))
    assert s.converged and sr.converged
    assert abs(s.computed_iterations - sr.computed_iterations) <= 1
    assert O.true_relres(A, b, x.numpy()) < 1e-9
    # FP32-inner refinement with the same preconditioner
    x2 = spk.DeviceVector.zeros(A.nrows)
    icfg = spk.IrConfig(1e-10, 20, 1e-4, 50, 500, spk.IR_GMRES, spk.ORTHO_CGS2)
    s2 = spk.ir_solve(spk.DeviceVector.from_numpy(b), x2, icfg, Ad, F)
    assert s2.converged and O.true_relres(A, b, x2.numpy()) < 1e-10


def _neumann_inverse(O, A):
    """An SPD approximate inverse on A's pattern: D^-1 - (1/2) D^-1 (A - D)
    D^-1 (first Neumann term, halved), for a symmetric diagonally dominant
    A."""
    n = A.nrows
    d = np.array([A.at(i, i) for i in range(n)])
    v = np.empty(A.nnz)
    for i in range(n):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            j = A.col_ind[k]
            v[k] = 1.0 / d[i] if i == j else -0.5 * A.vals[k] / (d[i] * d[j])
    return O.Csr(n, n, A.row_ptr.copy(), A.col_ind.copy(), v)


@pytest.mark.parametrize("max_it", [1, 5, 6, 11, 400])
def test_cg_with_explicit_approximate_inverse(O, spk, max_it):
    """CG with an explicit approximate inverse in the left slot (z = M r as
    an SpMV whose epilogue sums <r,z>; the x update rides in the direction
    pass) follows the oracle's conjugate_gradient with the same operator,
    also when stopped mid-way (restarts every 5 steps)."""
    A = O.aniso_q1_2d(20, 3, 1.5)
    M = _neumann_inverse(O, A)
    b = O.uniform_pm1(13, A.nrows)
    x0 = O.uniform_pm1(14, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    P = spk.spai(spk.CsrMatrix.from_oracle(M))
    xd = dv(spk, x0)
    restart = 5 if max_it < 400 else 1000
    cfg = spk.CGConfig(1e-10, restart, max_it, True, spk.F64)
    s = spk.conjugate_gradient(dv(spk, b), xd, cfg, Ad, P, spk.identity)
    xr, sr = O.cg(A, b, x0, tol=1e-10, restart_iterations=restart,
                  max_iterations=max_it, left=("csr", M), right=None)
    x = xd.numpy()
    if max_it < 400:
        assert not s.converged and s.computed_iterations == max_it
        assert np.allclose(x, xr, rtol=0, atol=1e-11 * np.abs(xr).max())
        assert np.allclose(s.iteration_residuals, sr.iteration_residuals,
                           rtol=1e-9)
    else:
        assert s.converged and sr.converged
        assert abs(s.computed_iterations - sr.computed_iterations) <= 1
        assert O.true_relres(A, b, x) < 1e-9


@pytest.mark.parametrize("max_it", [2, 4, 6, 9, 14])
@pytest.mark.parametrize("left,right", [(None, None), ("jacobi", None),
                                        (None, "ilu")])
def test_bicgstab_iterate_at_any_step(O, spk, max_it, left, right):
    """BiCGSTAB stopped by its matvec budget at any step (each step spends
    two; the half-step test on |s| and the step's |r| test both run) returns
    the oracle's iterate and residual history (bicgstab.hpp:176-324)."""
    A = O.convdiff_centered_2d(9, 8, 1.0, 0.4, 0.3, 2.0)
    b = O.uniform_pm1(13, A.nrows)
    x0 = O.uniform_pm1(14, A.nrows)
    s, x, sr, xr = _run(spk, O, "bicgstab", A, b, 1e-30, left, right,
                        max_it=max_it, x0=x0)
    assert not s.converged and s.computed_iterations == sr.computed_iterations
    assert np.allclose(x, xr, rtol=0, atol=1e-10 * np.abs(xr).max())
    assert np.allclose(s.iteration_residuals, sr.iteration_residuals, rtol=1e-8)
