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
