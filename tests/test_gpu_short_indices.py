"""16-bit column offsets for the row kernel (spk_csr_set_short_indices; no
reference counterpart): each index is read as an offset from its row
block's first row, so SpMVs are the same products in the same order --
equal bit for bit to the 32-bit reads and to the oracle's storage-order
multiply (sparse_blas.hpp:26-31) for one-lane rows -- in FP64, FP32 and
with FP16 values; stores whose band is too wide are rejected and keep
their 32-bit reads."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mult(spk, Ad, x):
    return spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()


@pytest.mark.parametrize("which", ["q1", "poisson3d", "convdiff"])
def test_short_indices_bit_exact(O, spk, which):
    A = {"q1": lambda: O.aniso_q1_2d(96, 5, 1.5),   # 9-point, band 97
         "poisson3d": lambda: O.poisson_3d(30, 29, 28),  # band 870
         "convdiff": lambda: O.convdiff_3d(24, 23, 22, 1.0, 0.5, 0.25, 7.0)}[which]()
    Ad = spk.CsrMatrix.from_oracle(A)
    nb, _ = Ad.plan_info()
    assert nb > 8  # several blocks, odd first entries among them
    x = O.uniform_pm1(3, A.nrows)
    x32 = x.astype(np.float32)
    ref64, ref32 = _mult(spk, Ad, x), _mult(spk, Ad, x32)
    Ad.set_half_values(True)
    ref16 = _mult(spk, Ad, x32)
    Ad.set_half_values(False)
    Ad.set_short_indices(True)
    y64 = _mult(spk, Ad, x)
    assert np.array_equal(y64, ref64)
    assert np.array_equal(y64, O.spmv(A, x))
    assert np.array_equal(_mult(spk, Ad, x32), ref32)
    Ad.set_half_values(True)
    assert np.array_equal(_mult(spk, Ad, x32), ref16)
    Ad.set_half_values(False)
    Ad.set_short_indices(False)
    assert np.array_equal(_mult(spk, Ad, x), ref64)


def test_short_indices_long_rows_and_edges(O, spk):
    """A row longer than a row block's entry limit (a block of its own,
    offsets from that row) beside short rows, with columns up to 32767
    either side of the block's first row."""
    rng = np.random.default_rng(2)
    n = 75000
    rows = []
    for i in range(n):
        if i == 40000:
            cols = np.arange(40000 - 32767, 40000 + 32767 + 1, 5)  # 13107 entries
            assert cols[-1] < n
        else:
            lo, hi = max(0, i - 3000), min(n, i + 3000)
            cols = np.unique(np.concatenate([[i], rng.integers(lo, hi, 4)]))
        rows.append(np.sort(cols))
    rp = np.concatenate([[0], np.cumsum([len(c) for c in rows])])
    ci = np.concatenate(rows).astype(np.int64)
    v = rng.uniform(-1, 1, len(ci))
    A = O.Csr(n, n, rp, ci, v)
    Ad = spk.CsrMatrix.from_host(n, n, rp, ci, v)
    x = O.uniform_pm1(4, n)
    ref = _mult(spk, Ad, x)
    Ad.set_short_indices(True)
    y = _mult(spk, Ad, x)
    assert np.array_equal(y, ref)
    short = np.array([len(c) for c in rows]) <= 64
    assert np.array_equal(y[short], O.spmv(A, x)[short])


def test_short_indices_reject_wide_band(O, spk):
    A = O.poisson_3d(200, 200, 2)  # band 40000
    Ad = spk.CsrMatrix.from_oracle(A)
    with pytest.raises(spk.InvalidArgument, match="32767"):
        Ad.set_short_indices(True)
    x = O.uniform_pm1(6, A.nrows)
    assert np.array_equal(_mult(spk, Ad, x), O.spmv(A, x))


def test_short_indices_with_panels_and_empty(O, spk):
    A = O.poisson_3d(20, 20, 20)
    Ad = spk.CsrMatrix.from_oracle(A)
    x = O.uniform_pm1(7, A.nrows)
    ref = _mult(spk, Ad, x)
    Ad.set_short_indices(True)
    Ad.set_panels(4)  # the panels keep their own indices
    y = _mult(spk, Ad, x)
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-14)
    Ad.set_panels(1)
    assert np.array_equal(_mult(spk, Ad, x), ref)
    E = spk.CsrMatrix.from_oracle(O.make_matrix(3, 3, [], [], []))
    E.set_short_indices(True)
    assert _mult(spk, E, np.ones(3)).tolist() == [0.0, 0.0, 0.0]


def test_short_indices_cg_same_iterates(O, spk):
    """Jacobi-PCG reads the same products, so the iterates agree bit for
    bit with the 32-bit reads."""
    A = O.poisson_3d(24, 22, 20)
    b = O.uniform_pm1(5, A.nrows)
    out = []
    for short in (False, True):
        Ad = spk.CsrMatrix.from_oracle(A)
        if short:
            Ad.set_short_indices(True)
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.conjugate_gradient(spk.DeviceVector.from_numpy(b), x,
                                   spk.CGConfig(1e-10, 5000, 5000), Ad,
                                   spk.jacobi(Ad))
        assert s.converged
        out.append((s.computed_iterations, x.numpy()))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])


def test_short_indices_learned_relcg(spk):
    """FP32 PCG with FP64 reliable updates and the learned inverse in FP16,
    both stores with short indices: the same 1e-8 FP64 true residual."""
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(16)]
    model = spk.SpaiModel(1)
    model.train(mats, epochs=4, learning_rate=2e-3, shuffle_seed=3)
    A = spk.aniso_q1_2d(128, 4242, 1.5)
    b = spk.uniform_pm1(42, A.nrows)
    S = model.predict(A).symmetric_part()
    S.set_half_values(True)
    cfg = spk.IrConfig(tolerance=1e-8, max_outer=200, inner_tolerance=1e-1,
                       restart=0, max_inner=20000, inner_method=spk.IR_CG_RELIABLE)
    its = []
    for short in (False, True):
        A.set_short_indices(short)
        S.set_short_indices(short)
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.ir_solve(b, x, cfg, A, spk.spai(S))
        assert s.converged and spk.true_relres(A, b, x) < 1e-8
        its.append(s.inner_iterations)
    assert its[0] == its[1]
