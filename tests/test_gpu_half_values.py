"""FP16 values for the FP32 SpMVs of a preconditioner store
(spk_csr_set_half_values; no reference counterpart): an FP32 SpMV reads the
FP32 values rounded to FP16, so it equals the oracle's FP32 storage-order
multiply (sparse_blas.hpp:26-31) on those rounded values bit for bit; FP64
SpMVs and the FP32 path after switching back are unchanged; the learned
FP32 PCG with an FP16 approximate inverse still reaches the FP64
true-residual target."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rounded(O, A):
    v = np.asarray(A.vals).astype(np.float32).astype(np.float16)
    return O.Csr(A.nrows, A.ncols, A.row_ptr, A.col_ind, v.astype(np.float64))


def test_half_values_spmv_bit_exact(O, spk):
    A = O.aniso_q1_2d(64, 3, 1.5)  # 9-point rows: one lane sums a row
    Ad = spk.CsrMatrix.from_oracle(A)
    x = O.uniform_pm1(5, A.nrows)
    x32 = x.astype(np.float32)
    y64 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()
    assert np.array_equal(y64, O.spmv(A, x))
    Ad.set_half_values(True)
    y16 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x32)).numpy()
    assert np.array_equal(y16, O.spmv(_rounded(O, A), x32))
    assert not np.array_equal(y16, O.spmv(A, x32))  # the rounding is visible
    assert np.array_equal(
        spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy(), y64)
    Ad.set_half_values(False)
    assert np.array_equal(
        spk.multiply(Ad, spk.DeviceVector.from_numpy(x32)).numpy(),
        O.spmv(A, x32))


def test_half_values_long_rows(O, spk):
    """Rows past a row block's entry limit (one block per row) and rows
    summed by two lanes read the FP16 values too (sums split over lanes:
    within FP32 rounding of the storage-order sum)."""
    rng = np.random.default_rng(4)
    n, nc = 300, 20000
    lens = np.where(np.arange(n) % 50 == 0, 6000, 20)
    cols = [np.sort(rng.choice(nc, size=k, replace=False)) for k in lens]
    rp = np.concatenate([[0], np.cumsum(lens)])
    ci = np.concatenate(cols).astype(np.int64)
    v = rng.uniform(-1, 1, len(ci))
    A = O.Csr(n, nc, rp, ci, v)
    Ad = spk.CsrMatrix.from_host(n, nc, rp, ci, v)
    Ad.set_half_values(True)
    x32 = O.uniform_pm1(9, nc).astype(np.float32)
    y = spk.multiply(Ad, spk.DeviceVector.from_numpy(x32)).numpy()
    ref = O.spmv(_rounded(O, A), x32).astype(np.float64)
    absrow = np.abs(_rounded(O, A).to_scipy()) @ np.abs(x32.astype(np.float64))
    assert np.all(np.abs(y - ref) <= 2e-6 * absrow)
    # against the FP32 values the FP16 rounding shows
    exact32 = O.spmv(A, x32).astype(np.float64)
    assert np.max(np.abs(y - exact32) / absrow) > 1e-5


def test_half_values_reject_overflow(spk):
    Ad = spk.CsrMatrix.from_host(2, 2, [0, 1, 2], [0, 1], [1.0, 1.0e6])
    with pytest.raises(spk.InvalidArgument, match="FP16"):
        Ad.set_half_values(True)
    x = spk.DeviceVector.from_numpy(np.ones(2, np.float32))
    assert np.array_equal(spk.multiply(Ad, x).numpy(), [1.0, 1.0e6])


def test_learned_relcg_with_half_inverse(spk):
    """FP32 PCG with FP64 reliable updates (spk_ir_solve) with the learned
    inverse's symmetric part in FP16: the same 1e-8 FP64 true residual,
    in about the FP32-valued inverse's iteration count."""
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(16)]
    model = spk.SpaiModel(1)
    model.train(mats, epochs=4, learning_rate=2e-3, shuffle_seed=3)
    A = spk.aniso_q1_2d(128, 4242, 1.5)
    b = spk.uniform_pm1(42, A.nrows)
    cfg = spk.IrConfig(tolerance=1e-8, max_outer=200, inner_tolerance=1e-1,
                       restart=0, max_inner=20000,
                       inner_method=spk.IR_CG_RELIABLE)
    S32 = model.predict(A).symmetric_part()
    S16 = model.predict(A).symmetric_part()
    S16.set_half_values(True)
    its = []
    for S in (S32, S16):
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.ir_solve(b, x, cfg, A, spk.spai(S))
        assert s.converged and spk.true_relres(A, b, x) < 1e-8
        its.append(s.inner_iterations)
    assert its[1] <= 1.1 * its[0]


def test_ilu_half_factors_in_fp32_refinement(spk):
    """The Jacobi-swept ILU(0) apply with FP16 strict factors: the FP32
    apply moves by about FP16's rounding, the FP64 apply not at all, and
    FP32-inner GMRES(50) under FP64 refinement still reaches 1e-10."""
    A = spk.convdiff_3d(40, 40, 40, 1.0, 1.35, 0.68, 0.34)
    F = spk.ilu0(A)
    F.jacobi_sweeps = 3
    r = spk.uniform_pm1(3, A.nrows)
    r32 = r.astype(np.float32)

    def apply(v):
        z = spk.DeviceVector.zeros(A.nrows, v.dtype)
        F.apply(v, z)
        return z.numpy()

    z64, z32 = apply(r), apply(r32)
    F.half_factors = True
    assert F.half_factors
    h32 = apply(r32)
    assert np.array_equal(apply(r), z64)
    d = np.linalg.norm(h32.astype(np.float64) - z32) / np.linalg.norm(z32)
    assert 1e-6 < d < 1e-2
    b = spk.uniform_pm1(42, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.IrConfig(1e-10, 20, 1e-4, 50, 2000, spk.IR_GMRES, spk.ORTHO_CGS2)
    s = spk.ir_solve(b, x, cfg, A, F)
    assert s.converged and spk.true_relres(A, b, x) < 1e-10
    F.half_factors = False
    assert not F.half_factors
    assert np.array_equal(apply(r32), z32)


def test_half_values_empty_store(O, spk):
    E = spk.CsrMatrix.from_oracle(O.make_matrix(3, 3, [], [], []))
    E.set_half_values(True)
    x = spk.DeviceVector.from_numpy(np.array([1., 2., 3.], np.float32))
    assert spk.multiply(E, x).numpy().tolist() == [0., 0., 0.]
