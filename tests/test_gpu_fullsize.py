"""Parity at the benchmark configurations' full sizes.  Where the oracle is
fast enough at full size the device result is compared with it bit for bit
(the 192^3 Poisson SpMV, the 128^3 convection-diffusion ILU(0) factors, the
1024^2 anisotropic Q1 matrix and its network features); otherwise a
size-independent property is checked (the power-law matrix's exact row sums
and symmetry, the batch's per-system true residuals).  Every solve is judged
on its FP64 true residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dv(spk, a):
    return spk.DeviceVector.from_numpy(np.ascontiguousarray(a))


def test_poisson_192_spmv_bit_exact_and_pcg(O, spk):
    print("full size: poisson_3d(192) SpMV and Jacobi-PCG", flush=True)
    R = O.poisson_3d(192)
    A = spk.poisson_3d(192, 192, 192)
    rp, ci, v = A.download()
    assert np.array_equal(rp, R.row_ptr) and np.array_equal(ci, R.col_ind)
    assert np.array_equal(v, R.vals)
    x = O.uniform_pm1(11, A.nrows)
    y = spk.DeviceVector.zeros(A.nrows)
    A.spmv(_dv(spk, x), y)
    assert np.array_equal(y.numpy(), O.spmv(R, x))  # storage-order sums
    b = spk.uniform_pm1(42, A.nrows)
    xs = spk.DeviceVector.zeros(A.nrows)
    s = spk.conjugate_gradient(b, xs, spk.CGConfig(1e-8, 20000, 20000), A,
                               spk.jacobi(A))
    assert s.converged and 640 <= s.computed_iterations <= 648
    assert spk.true_relres(A, b, xs) < 1e-8


@pytest.mark.parametrize("max_it", [5, 6])
def test_poisson_192_pcg_iterate_at_step(O, spk, max_it):
    """Above 4096 SpMV row blocks Jacobi-PCG keeps the SpMV's finisher and
    hands the r update's partials to the x / p pass: a solve stopped at a
    restart step and one past it return the oracle's iterate and residual
    history (conjugate_gradient.hpp:208-292)."""
    print(f"full size: poisson_3d(192) Jacobi-PCG stopped after {max_it}", flush=True)
    R = O.poisson_3d(192)
    A = spk.CsrMatrix.from_oracle(R)
    b = O.uniform_pm1(11, R.nrows)
    x0 = O.uniform_pm1(12, R.nrows)
    xd = _dv(spk, x0)
    cfg = spk.CGConfig(1e-30, 5, max_it, True, spk.F64)
    s = spk.conjugate_gradient(_dv(spk, b), xd, cfg, A, spk.jacobi(A), spk.identity)
    xr, sr = O.cg(R, b, x0, tol=1e-30, restart_iterations=5,
                  max_iterations=max_it, left="jacobi", right=None)
    assert not s.converged and s.computed_iterations == max_it
    assert np.allclose(xd.numpy(), xr, rtol=0, atol=1e-11 * np.abs(xr).max())
    assert np.allclose(s.iteration_residuals, sr.iteration_residuals, rtol=1e-9)


def test_powerlaw_4m_row_sums_symmetry_and_pcg(spk):
    """Off-diagonals in [-1, 0] and diagonal = sum |off| + 1 (matgen.hpp:
    657-658): every row of A sums to 1 up to rounding; A is symmetric, so
    <y, A x> = <A y, x>."""
    print("full size: powerlaw_spd(4194304) properties and PCG", flush=True)
    A = spk.powerlaw_spd(4194304, 42)
    n = A.nrows
    ones = _dv(spk, np.ones(n))
    y = spk.DeviceVector.zeros(n)
    A.spmv(ones, y)
    assert np.abs(y.numpy() - 1.0).max() < 1e-9
    rng = np.random.default_rng(3)
    xa, ya = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    ax, ay = spk.DeviceVector.zeros(n), spk.DeviceVector.zeros(n)
    A.spmv(_dv(spk, xa), ax)
    A.spmv(_dv(spk, ya), ay)
    l, r = np.dot(ya, ax.numpy()), np.dot(ay.numpy(), xa)
    assert abs(l - r) <= 1e-9 * (abs(l) + abs(r))
    b = spk.uniform_pm1(42, n)
    xs = spk.DeviceVector.zeros(n)
    s = spk.conjugate_gradient(b, xs, spk.CGConfig(1e-8, 20000, 20000), A,
                               spk.jacobi(A))
    assert s.converged and spk.true_relres(A, b, xs) < 1e-8


def test_convdiff_128_ilu_bit_exact_and_refinement(O, spk):
    print("full size: convdiff_3d(128) ILU(0) factors and refinement", flush=True)
    pe = 63.272
    h = 1.0 / 129
    v = np.array([1.0, 0.5, 0.25])
    c = pe * h * v / np.linalg.norm(v)
    R = O.convdiff_3d(128, 128, 128, 1.0, *c)
    A = spk.CsrMatrix.from_oracle(R)
    F = spk.ilu0(A)
    (lrp, lci, lv), (urp, uci, uv) = F.factors()
    Lr, Ur = O.ilu0(R)
    assert np.array_equal(lrp, Lr.row_ptr) and np.array_equal(lci, Lr.col_ind)
    assert np.array_equal(urp, Ur.row_ptr) and np.array_equal(uci, Ur.col_ind)
    assert np.array_equal(lv, Lr.vals) and np.array_equal(uv, Ur.vals)
    b = spk.uniform_pm1(42, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    F.jacobi_sweeps = 3
    cfg = spk.IrConfig(1e-10, 20, 1e-4, 50, 2000, spk.IR_GMRES, spk.ORTHO_CGS2)
    s = spk.ir_solve(b, x, cfg, A, F)
    assert s.converged and spk.true_relres(A, b, x) < 1e-10


def test_aniso_q1_1024_bit_exact_features_and_learned_pcg(O, spk):
    print("full size: aniso_q1_2d(1024) matrix, features, learned PCG", flush=True)
    R = O.aniso_q1_2d(1024, 4242, 1.5)
    A = spk.aniso_q1_2d(1024, 4242, 1.5)
    rp, ci, v = A.download()
    assert np.array_equal(rp, R.row_ptr) and np.array_equal(ci, R.col_ind)
    assert np.array_equal(v, R.vals)
    X = spk.spai_features(A)
    assert np.array_equal(X, O.spai_features(R))
    mats = [spk.aniso_q1_2d(256, s, 1.5) for s in range(32)]
    model = spk.SpaiModel(1)
    model.train(mats, epochs=2, learning_rate=2e-3, shuffle_seed=7)
    Ms = spk.spai(model.predict(A).symmetric_part())
    b = spk.uniform_pm1(42, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.conjugate_gradient(b, x, spk.CGConfig(1e-8, 50000, 50000), A, Ms)
    assert s.converged and spk.true_relres(A, b, x) < 1e-8
    assert s.computed_iterations < 9444  # Jacobi-PCG's count on this matrix


def test_batched_4096_every_system(spk):
    print("full size: 4096 batched systems", flush=True)
    B = spk.BatchedSystems.diffusion(4096, 42, 32, 256)
    b = spk.uniform_pm1(5, B.total_rows)
    x = spk.DeviceVector.zeros(B.total_rows)
    s = B.solve(spk.BatchConfig(1e-8, 1e-2, 40, 20000, spk.F32), b, x)
    assert s["converged"] == 4096 and s["max_relres"] < 1e-8
