"""Mixed-precision iterative refinement on the GPU: FP32 inner Krylov
(FP64 dots and scalars) with an FP64 true-residual outer loop reaches FP64
accuracy that an FP32-only solve cannot, and agrees with the oracle's FP64
solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def convdiff_problem(O, nx, ny, nz, pe):
    v = np.array([1.0, 0.5, 0.25])
    h = 1.0 / (nx + 1)
    c = pe * h * v / np.linalg.norm(v)
    return O.convdiff_3d(nx, ny, nz, 1.0, *c), c


@pytest.mark.parametrize("method", ["gmres", "fgmres", "bicgstab"])
def test_ir_reaches_fp64_accuracy(O, spk, method):
    A, c = convdiff_problem(O, 24, 20, 16, 60.0)
    b = O.uniform_pm1(42, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    F = spk.ilu0(Ad)
    x = spk.DeviceVector.zeros(A.nrows)
    m = {"gmres": spk.IR_GMRES, "fgmres": spk.IR_FGMRES,
         "bicgstab": spk.IR_BICGSTAB}[method]
    cfg = spk.IrConfig(tolerance=1e-10, max_outer=20, inner_tolerance=1e-4,
                       restart=50, max_inner=400, inner_method=m)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, cfg, Ad, F)
    assert s.converged
    xh = x.numpy()
    rel = O.true_relres(A, b, xh)
    assert rel < 1e-10
    assert abs(rel - s.relres) <= 1e-3 * rel + 1e-16
    # the reported history is the FP64 true residual, decreasing
    h = np.array(s.relres_history)
    assert np.all(h[1:] < h[:-1])
    # agrees with the oracle's FP64 GMRES + ILU(0) solution
    xr, sr = O.gmres(A, b, tol=1e-12, restart_dimension=50,
                     max_iterations=2000, right="ilu")
    assert np.linalg.norm(xh - xr) <= 1e-8 * np.linalg.norm(xr)


def test_ir_cg_jacobi(O, spk):
    A = O.poisson_3d(20, 18, 16)
    b = O.uniform_pm1(7, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.IrConfig(tolerance=1e-11, max_outer=20, inner_tolerance=1e-4,
                       restart=100000, max_inner=2000, inner_method=spk.IR_CG)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, cfg, Ad, spk.jacobi(Ad))
    assert s.converged and O.true_relres(A, b, x.numpy()) < 1e-11


def test_fp32_alone_stagnates(O, spk):
    """Without refinement an FP32 Krylov solve stalls far above 1e-10."""
    A, c = convdiff_problem(O, 24, 20, 16, 60.0)
    b = O.uniform_pm1(42, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    F = spk.ilu0(Ad)
    x32 = spk.DeviceVector.zeros(A.nrows, np.float32)
    cfg = spk.GmresConfig(1e-10, 50, 300, False, spk.F32)
    spk.gmres(spk.DeviceVector.from_numpy(b.astype(np.float32)), x32, cfg, Ad,
              spk.identity, F)
    assert O.true_relres(A, b, x32.numpy().astype(np.float64)) > 1e-9


def test_ir_zero_rhs_and_exact_start(O, spk):
    A, c = convdiff_problem(O, 10, 9, 8, 30.0)
    Ad = spk.CsrMatrix.from_oracle(A)
    x = spk.DeviceVector.from_numpy(np.ones(A.nrows))
    s = spk.ir_solve(spk.DeviceVector.zeros(A.nrows), x, spk.IrConfig(), Ad)
    assert s.converged and np.all(x.numpy() == 0.0)
    xt = O.uniform_pm1(3, A.nrows)
    b = O.spmv(A, xt)
    x = spk.DeviceVector.from_numpy(xt)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, spk.IrConfig(), Ad)
    assert s.converged and s.outer_iterations == 0


def _relcg(spk, Ad, b, M, tol=1e-10, delta=1e-2, max_outer=50, max_inner=5000):
    x = spk.DeviceVector.zeros(Ad.nrows)
    cfg = spk.IrConfig(tolerance=tol, max_outer=max_outer, inner_tolerance=delta,
                       restart=0, max_inner=max_inner,
                       inner_method=spk.IR_CG_RELIABLE)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(b), x, cfg, Ad, M)
    return s, x.numpy()


@pytest.mark.parametrize("prec", ["jacobi", "identity", "spai", "ilu"])
def test_cg_reliable_updates(O, spk, prec):
    """FP32 PCG with FP64 reliable updates reaches the FP64 true residual
    that FP32 alone cannot, in about the FP64 PCG's iteration count (the
    search direction survives each update), and agrees with the oracle's
    FP64 CG solution."""
    A = O.poisson_3d(22, 20, 18)
    b = O.uniform_pm1(5, A.nrows)
    Ad = spk.CsrMatrix.from_oracle(A)
    left = {"jacobi": "jacobi", "identity": None, "spai": "jacobi",
            "ilu": "ilu"}[prec]
    M = {"jacobi": lambda: spk.jacobi(Ad), "identity": lambda: None,
         # an untrained network predicts the Jacobi inverse exactly
         "spai": lambda: spk.spai(spk.SpaiModel(1).predict(Ad).symmetric_part()),
         "ilu": lambda: spk.ilu0(Ad)}[prec]()
    s, x = _relcg(spk, Ad, b, M, tol=1e-11)
    assert s.converged and s.outer_iterations >= 2
    rel = O.true_relres(A, b, x)
    assert rel < 1e-11 and abs(rel - s.relres) <= 1e-3 * rel
    h = np.array(s.relres_history)
    assert len(h) == s.outer_iterations + 1 and np.all(h[1:] < h[:-1])
    xr, sr = O.cg(A, b, tol=1e-13, restart_iterations=100000,
                  max_iterations=5000, left=left)
    assert sr.converged
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
    _, s64 = O.cg(A, b, tol=1e-11, restart_iterations=100000,
                  max_iterations=5000, left=left)
    assert s.inner_iterations <= 1.15 * s64.computed_iterations + 5, (
        s.inner_iterations, s64.computed_iterations)


def test_cg_reliable_budget_and_exact_start(O, spk):
    A = O.poisson_3d(12, 11, 10)
    Ad = spk.CsrMatrix.from_oracle(A)
    b = O.uniform_pm1(6, A.nrows)
    s, x = _relcg(spk, Ad, b, spk.jacobi(Ad), tol=1e-12, delta=1e-9,
                  max_inner=7)
    assert not s.converged and s.inner_iterations == 7
    # the partial correction is kept: the residual went down
    assert O.true_relres(A, b, x) < 1.0
    xt = O.uniform_pm1(3, A.nrows)
    bt = O.spmv(A, xt)
    xd = spk.DeviceVector.from_numpy(xt)
    cfg = spk.IrConfig(tolerance=1e-10, inner_method=spk.IR_CG_RELIABLE)
    s = spk.ir_solve(spk.DeviceVector.from_numpy(bt), xd, cfg, Ad, spk.jacobi(Ad))
    assert s.converged and s.outer_iterations == 0 and s.inner_iterations == 0
    with pytest.raises(spk.InvalidArgument):
        spk.ir_solve(spk.DeviceVector.from_numpy(bt), xd,
                     spk.IrConfig(inner_method=7), Ad)


@pytest.mark.parametrize("prec", ["jacobi", "learned"])
def test_cg_reliable_ill_conditioned(O, spk, prec):
    """The FP32 recurrence with FP64 reliable updates on an ill-conditioned
    anisotropic Q1 matrix (conductivity lognormal(0, 1.5), anisotropy up to
    1e3; the small problems above are well conditioned): it reaches the
    FP64 true-residual tolerance, with the learned inverse's symmetric part
    in about the FP64 PCG's iteration count."""
    A = spk.aniso_q1_2d(128, 4242, 1.5)
    n = A.nrows
    b = spk.uniform_pm1(42, n)
    if prec == "jacobi":
        M = spk.jacobi(A)
    else:
        mats = [spk.aniso_q1_2d(128, s, 1.5) for s in range(32)]
        model = spk.SpaiModel(1)
        model.train(mats, epochs=8, learning_rate=2e-3, shuffle_seed=7)
        M = spk.spai(model.predict(A).symmetric_part())
    x64 = spk.DeviceVector.zeros(n)
    s64 = spk.conjugate_gradient(b, x64, spk.CGConfig(1e-10, 20000, 20000), A, M)
    assert s64.converged
    x = spk.DeviceVector.zeros(n)
    cfg = spk.IrConfig(tolerance=1e-10, max_outer=100, inner_tolerance=1e-1,
                       restart=0, max_inner=40000, inner_method=spk.IR_CG_RELIABLE)
    s = spk.ir_solve(b, x, cfg, A, M)
    assert s.converged and spk.true_relres(A, b, x) < 1e-10
    if prec == "learned":
        assert s.inner_iterations <= 1.2 * s64.computed_iterations + 10, (
            s.inner_iterations, s64.computed_iterations)
