"""GPU parity of the device store, SpMV, diagonal extraction and generators
against the oracle (sparse_blas.hpp:17-34, unary.hpp:18-39,
Compressed_row_sparsity_Impl.cpp:10-40, matgen.hpp)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_csr(O, rng, nrows, ncols, lengths):
    rows, cols = [], []
    for r, L in enumerate(lengths):
        c = rng.choice(ncols, size=min(L, ncols), replace=False)
        rows += [r] * len(c)
        cols += list(c)
    vals = rng.uniform(-1, 1, len(rows))
    return O.make_matrix(nrows, ncols, rows, cols, vals)


def _spmv_dev(spk, A, x, dtype=np.float64):
    Ad = spk.CsrMatrix.from_oracle(A)
    return spk.multiply(Ad, spk.DeviceVector.from_numpy(x.astype(dtype))).numpy()


def test_spmv_known_3x3(O, spk):
    # sparse_blas_test.cpp:36-54
    A = O.make_matrix(3, 3, [0, 0, 1, 2, 2], [0, 2, 1, 0, 2],
                      [2., 3., 4., 5., 6.])
    y = _spmv_dev(spk, A, np.array([1., 2., 3.]))
    assert y.tolist() == [11.0, 8.0, 23.0]


def test_spmv_empty_rectangular_identity(O, spk):
    # sparse_blas_test.cpp:56-95
    E = O.make_matrix(3, 3, [], [], [])
    assert _spmv_dev(spk, E, np.array([1., 2., 3.])).tolist() == [0., 0., 0.]
    R = O.make_matrix(2, 3, [0, 0, 1], [0, 2, 1], [1., 2., 3.])
    assert _spmv_dev(spk, R, np.array([1., 2., 3.])).tolist() == [7., 6.]
    I = O.make_matrix(3, 3, [0, 1, 2], [0, 1, 2], [1., 1., 1.])
    assert _spmv_dev(spk, I, np.array([4., 5., 6.])).tolist() == [4., 5., 6.]


def test_spmv_short_rows_bit_exact(O, spk):
    """Rows summed by one lane reproduce the reference's sequential sum
    bit for bit."""
    A = O.poisson_2d(96, 80, 0.25)
    x = O.uniform_pm1(3, A.ncols)
    assert np.array_equal(_spmv_dev(spk, A, x), O.spmv(A, x))
    B = O.convdiff_3d(20, 17, 13, 1.0, 0.3, 0.2, 0.1)
    xb = O.uniform_pm1(4, B.ncols)
    assert np.array_equal(_spmv_dev(spk, B, xb), O.spmv(B, xb))


@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_mixed_row_lengths(O, spk, seed):
    """Empty rows, short rows, medium rows (several lanes per row) and rows
    longer than a block's nonzero budget (block-wide reduction)."""
    rng = np.random.default_rng(seed)
    n = 3000
    lengths = rng.choice([0, 1, 3, 7, 40, 300, 1500, 4500, 6000], size=n,
                         p=[.05, .15, .3, .3, .1, .06, .02, .01, .01])
    A = _random_csr(O, rng, n, 7000, lengths)
    x = rng.uniform(-1, 1, A.ncols)
    y = _spmv_dev(spk, A, x)
    y_ref = O.spmv(A, x)
    absrow = np.abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-12 * absrow + 1e-300)
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref)
    Ad = spk.CsrMatrix.from_oracle(A)
    nb, nl = Ad.plan_info()
    assert nl == int(np.sum(lengths > 4096))


def test_spmv_fp32(O, spk):
    rng = np.random.default_rng(5)
    lengths = rng.integers(0, 60, 5000)
    A = _random_csr(O, rng, 5000, 5000, lengths)
    x = rng.uniform(-1, 1, A.ncols).astype(np.float32)
    y = _spmv_dev(spk, A, x, np.float32)
    y_ref = O.spmv(A, x.astype(np.float64))
    absrow = np.abs(A.to_scipy()) @ np.abs(x.astype(np.float64))
    assert np.all(np.abs(y - y_ref) <= 2e-6 * absrow + 1e-30)


def test_store_round_trip(O, spk):
    A = O.convection_diffusion_2d(30, 20, 1.0, 0.4, 0.1)
    Ad = spk.CsrMatrix.from_oracle(A)
    rp, ci, v = Ad.download()
    assert np.array_equal(rp, A.row_ptr)
    assert np.array_equal(ci, A.col_ind)
    assert np.array_equal(v, A.vals)


def test_store_rejects_bad_pattern(spk):
    with pytest.raises(spk.InvalidArgument):
        spk.CsrMatrix.from_host(2, 2, [0, 1, 2], [0, 5], [1.0, 1.0])
    with pytest.raises(spk.InvalidArgument):
        spk.CsrMatrix.from_host(2, 2, [0, 2, 1], [0, 1], [1.0, 1.0])
    # Shape(r, c) with r <= 1 or c <= 1 is a logic_error (Shape.cpp:8-10)
    with pytest.raises(spk.LogicError):
        spk.CsrMatrix.from_host(1, 3, [0, 1], [0], [1.0])


def test_extract_diagonal(O, spk):
    # unary_test.cpp:35-95: known values, rectangular, missing entries
    A = O.make_matrix(3, 4, [0, 0, 1, 2, 2], [0, 3, 2, 0, 2],
                      [5., 1., 7., 2., 9.])
    d = spk.extract_diagonal(spk.CsrMatrix.from_oracle(A)).numpy()
    assert d.tolist() == O.extract_diagonal(A).tolist() == [5.0, 0.0, 9.0]


def test_generators_bit_exact(O, spk):
    cases = [
        (spk.poisson_2d(33, 17, 0.0), O.poisson_2d(33, 17, 0.0)),
        (spk.poisson_2d(16, 16, 5.0), O.poisson_2d(16, 16, 5.0)),
        (spk.poisson_3d(9, 7, 5), O.poisson_3d(9, 7, 5)),
        (spk.convdiff_3d(8, 6, 5, 1.0, 0.37, 0.18, 0.09),
         O.convdiff_3d(8, 6, 5, 1.0, 0.37, 0.18, 0.09)),
    ]
    for Ad, A in cases:
        rp, ci, v = Ad.download()
        assert np.array_equal(rp, A.row_ptr)
        assert np.array_equal(ci, A.col_ind)
        assert np.array_equal(v, A.vals)


def test_uniform_rhs_bit_exact(O, spk):
    assert np.array_equal(spk.uniform_pm1(42, 10001).numpy(),
                          O.uniform_pm1(42, 10001))


def test_true_relres(O, spk):
    A = O.poisson_2d(40, 40, 1.0)
    b = O.uniform_pm1(1, A.nrows)
    x = O.uniform_pm1(2, A.nrows)
    r = spk.true_relres(spk.CsrMatrix.from_oracle(A),
                        spk.DeviceVector.from_numpy(b),
                        spk.DeviceVector.from_numpy(x))
    assert abs(r - O.true_relres(A, b, x)) <= 1e-13 * r


@pytest.mark.parametrize("n,seed", [(3000, 1), (20000, 7)])
def test_powerlaw_generator_bit_exact(O, spk, n, seed):
    """Device sort/unique/scan build of the power-law SPD matrix equals the
    oracle's sequential build bit for bit."""
    A = O.powerlaw_spd(n, seed)
    rp, ci, v = spk.powerlaw_spd(n, seed).download()
    assert np.array_equal(rp, A.row_ptr)
    assert np.array_equal(ci, A.col_ind)
    assert np.array_equal(v, A.vals)


def test_powerlaw_spmv(O, spk):
    """Medium and long power-law rows (several lanes per row, block-wide
    reduction) against the oracle."""
    A = O.powerlaw_spd(30000, 3)
    x = O.uniform_pm1(5, A.nrows)
    y = spk.multiply(spk.powerlaw_spd(30000, 3), spk.DeviceVector.from_numpy(x)).numpy()
    y_ref = O.spmv(A, x)
    absrow = np.abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-12 * absrow)


@pytest.mark.parametrize("panels", [2, 4, 7])
def test_panelled_spmv_bit_exact_on_stencils(O, spk, panels):
    """Column panels carry each row's running sum in column order, so a
    one-lane row keeps the storage-order sum: bit-identical to the oracle."""
    for A in (O.poisson_3d(11, 9, 13), O.poisson_2d(40, 31, 0.5)):
        Ad = spk.CsrMatrix.from_oracle(A)
        Ad.set_panels(panels)
        assert Ad.panels == panels
        x = O.uniform_pm1(3, A.nrows)
        y = spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()
        assert np.array_equal(y.view(np.uint64), O.spmv(A, x).view(np.uint64))
    Ad.set_panels(1)
    assert Ad.panels == 1


def test_panelled_powerlaw_and_solvers(O, spk):
    A = O.powerlaw_spd(30000, 3)
    x = O.uniform_pm1(5, A.nrows)
    Ad = spk.powerlaw_spd(30000, 3)
    y1 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()
    Ad.set_panels(8)
    y8 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()
    absrow = np.abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(y8 - O.spmv(A, x)) <= 1e-12 * absrow)
    assert np.all(np.abs(y8 - y1) <= 1e-12 * absrow)
    # FP32 SpMV through the panels
    y32 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x.astype(np.float32)))
    assert np.all(np.abs(y32.numpy() - y1) <= 1e-5 * absrow)
    # a PCG solve on the panelled store (fused epilogues run on the last panel)
    b = spk.uniform_pm1(9, A.nrows)
    its = []
    for P in (1, 8):
        Ad.set_panels(P)
        xs = spk.DeviceVector.zeros(A.nrows)
        s = spk.conjugate_gradient(b, xs, spk.CGConfig(1e-10, 1000, 1000), Ad,
                                   spk.jacobi(Ad))
        assert s.converged and spk.true_relres(Ad, b, xs) < 1e-9
        its.append(s.computed_iterations)
    assert abs(its[0] - its[1]) <= 1


def test_tune_layout(spk):
    Ad = spk.powerlaw_spd(300000, 4)
    chosen, tried = Ad.tune_layout()
    assert tried[0][0] == 1 and chosen in [p for p, _ in tried]
    assert Ad.panels == chosen
    assert dict(tried)[chosen] == min(us for _, us in tried)


def test_symmetric_part_bit_exact(O, spk):
    """(M + M^T)/2 on the device: the oracle's values bit for bit; a
    structurally nonsymmetric pattern is rejected."""
    for A in (O.convdiff_centered_2d(17, 13, 1.0, 0.3, 0.2, 1.0),
              O.aniso_q1_2d(24, 5, 1.5)):
        S = spk.CsrMatrix.from_oracle(A).symmetric_part()
        rp, ci, v = S.download()
        R = O.symmetric_part(A)
        assert np.array_equal(rp, R.row_ptr) and np.array_equal(ci, R.col_ind)
        assert np.array_equal(v, R.vals)
    M = O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 1, 2], [1., 2., 3., 4.])
    with pytest.raises(spk.InvalidArgument):
        spk.CsrMatrix.from_oracle(M).symmetric_part()


def test_gather_rate_ceiling(spk):
    """The gather-throughput measurement: a window that fits the L1 gathers
    faster than one spread over 32 MB; arguments are checked."""
    small, sb = spk.gather_rate(2048, 8 << 20, 2)
    large, lb = spk.gather_rate(4 << 20, 8 << 20, 2)
    assert small > large > 1e9
    assert abs(sb / small - 12.0) < 1e-9
    with pytest.raises(spk.InvalidArgument):
        spk.gather_rate(0, 4096, 1)
    with pytest.raises(spk.InvalidArgument):
        spk.gather_rate(1024, 4095, 1)


@pytest.mark.parametrize("case", ["tridiag4000", "poisson3d", "rect"])
def test_spmv_odd_block_starts(O, spk, case):
    """Row blocks whose first entry sits at an odd index: the kernel loads
    (index, value) pairs from the even start below it, so the pair's first
    element is the previous block's last entry, gathered but never summed.
    Bit-exact against the oracle in FP64 and FP32 over matrices with many
    odd block starts, the last block included, and a rectangular one."""
    A = {"tridiag4000": lambda: O.tridiagonal_matrix(4000, -1.0, 4.0, -1.0),
         "poisson3d": lambda: O.poisson_3d(20, 18, 16),
         "rect": lambda: O.make_matrix(3001, 7, list(range(3001)),
                                       [(3 * i) % 7 for i in range(3001)],
                                       [1.0 + i % 5 for i in range(3001)])}[case]()
    Ad = spk.CsrMatrix.from_oracle(A)
    x = O.uniform_pm1(21, A.ncols)
    y = spk.DeviceVector.zeros(A.nrows)
    Ad.spmv(spk.DeviceVector.from_numpy(x), y)
    assert np.array_equal(y.numpy(), O.spmv(A, x))
    x32 = x.astype(np.float32)
    y32 = spk.DeviceVector.zeros(A.nrows, np.float32)
    Ad.spmv(spk.DeviceVector.from_numpy(x32), y32)
    assert np.array_equal(y32.numpy(), O.spmv(A, x32))


def test_panels_with_few_and_no_rows(O, spk):
    """Inner column panels keep only their non-empty rows: here panel 1
    holds entries of every 37th row only and panel 2 none at all.  The
    running sums still follow column order, so one-lane rows stay
    bit-identical to the oracle, in FP64 and FP32."""
    n = 4000
    rows, cols, vals = [], [], []
    for i in range(n):
        rows += [i, i]
        cols += [i % 1000, 3000 + (7 * i) % 1000]
        vals += [1.0 + i % 5, -0.5 - i % 3]
        if i % 37 == 0:
            rows.append(i)
            cols.append(1000 + (3 * i) % 1000)
            vals.append(0.25 + i % 7)
    A = O.make_matrix(n, n, rows, cols, vals)
    Ad = spk.CsrMatrix.from_oracle(A)
    Ad.set_panels(4)
    assert Ad.panels == 4
    x = O.uniform_pm1(23, n)
    y = spk.DeviceVector.zeros(n)
    Ad.spmv(spk.DeviceVector.from_numpy(x), y)
    assert np.array_equal(y.numpy(), O.spmv(A, x))
    x32 = x.astype(np.float32)
    y32 = spk.DeviceVector.zeros(n, np.float32)
    Ad.spmv(spk.DeviceVector.from_numpy(x32), y32)
    assert np.array_equal(y32.numpy(), O.spmv(A, x32))


def test_spmv_dot_matches_oracle(O, spk):
    """spk_csr_spmv_dot: y = A x bit for bit as multiply (sparse_blas.hpp:
    17-34) and <x, y> (compute_step_size, conjugate_gradient.hpp:232-239)
    within FP64 rounding of the oracle's left-to-right sum; with column
    panels too; a non-square store is rejected."""
    R = O.poisson_3d(30, 28, 26)
    A = spk.CsrMatrix.from_oracle(R)
    x = O.uniform_pm1(3, R.nrows)
    y_ref = O.spmv(R, x)
    d_ref = float(np.dot(x, y_ref))
    for panels in (1, 3):
        A.set_panels(panels)
        y = spk.DeviceVector.zeros(R.nrows)
        r = spk.DeviceVector.zeros(1)
        A.spmv_dot(spk.DeviceVector.from_numpy(x), y, r)
        assert np.array_equal(y.numpy(), y_ref)
        assert abs(r.numpy()[0] - d_ref) <= 1e-12 * np.abs(x * y_ref).sum()
    B = spk.CsrMatrix.from_host(3, 4, [0, 1, 2, 3], [0, 1, 3], [1.0, 2.0, 3.0])
    with pytest.raises(spk.InvalidArgument):
        B.spmv_dot(spk.DeviceVector.zeros(4), spk.DeviceVector.zeros(3),
                   spk.DeviceVector.zeros(1))


@pytest.mark.parametrize("slots,window", [(16, 8192), (8, 8192), (16, 4096),
                                          (8, 256)])
def test_window_tiles_bit_exact(O, spk, slots, window):
    """Window tiles sum every tiled row in column order with separately
    rounded products and sums: the oracle's storage-order multiply
    (sparse_blas.hpp:26-31) bit for bit, FP64 and FP32.  Rows longer than
    the tiles' limit (at least 256 entries) run in the row kernel, which
    splits a long row over lanes: within 1e-12 of |A| |x| there.  Cases: the
    power-law matrix (rows across all columns, long rows, 2.5 row groups), a
    rectangular random matrix, and one x window holding many staging chunks
    per group; the epilogue paths (<x, A x>, CG) run on it too."""
    A = O.powerlaw_spd(40000, 6)
    lens = np.diff(A.row_ptr)
    tiled = lens <= 256  # tiled for every window size
    assert (~tiled).sum() > 100
    Ad = spk.CsrMatrix.from_oracle(A)
    Ad.set_window_tiles(slots, window)
    assert Ad.panels == -(slots * 65536 + window)
    x = O.uniform_pm1(5, A.nrows)
    y = spk.multiply(Ad, spk.DeviceVector.from_numpy(x)).numpy()
    y_ref = O.spmv(A, x)
    assert np.array_equal(y[tiled], y_ref[tiled])
    absrow = np.abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-12 * absrow)
    x32 = x.astype(np.float32)
    y32 = spk.multiply(Ad, spk.DeviceVector.from_numpy(x32)).numpy()
    y32_ref = O.spmv(A, x32)
    assert np.array_equal(y32[tiled], y32_ref[tiled])
    assert np.all(np.abs(y32.astype(np.float64) - y_ref) <= 2e-6 * absrow)
    yd = spk.DeviceVector.zeros(A.nrows)
    r = spk.DeviceVector.zeros(1)
    Ad.spmv_dot(spk.DeviceVector.from_numpy(x), yd, r)
    assert np.array_equal(yd.numpy(), y)
    assert abs(r.numpy()[0] - np.dot(x, y)) <= 1e-12 * np.abs(x * y).sum()
    rng = np.random.default_rng(slots + window)
    for nr, nc, lo, hi in ((20000, 70001, 0, 60), (20000, 8192, 8, 9)):
        # the second: one window, 8 entries a row, so each group's window
        # holds 131072 entries, 19 staging chunks
        cols = [np.sort(rng.choice(nc, size=rng.integers(lo, hi), replace=False))
                for _ in range(nr)]
        rp = np.concatenate([[0], np.cumsum([len(c) for c in cols])])
        ci = np.concatenate(cols).astype(np.int64)
        v = rng.uniform(-1, 1, len(ci))
        R = O.Csr(nr, nc, rp, ci, v)
        Rd = spk.CsrMatrix.from_host(nr, nc, rp, ci, v)
        Rd.set_window_tiles(slots, window)
        xr = O.uniform_pm1(8, nc)
        assert np.array_equal(
            spk.multiply(Rd, spk.DeviceVector.from_numpy(xr)).numpy(),
            O.spmv(R, xr))


def test_window_tiles_in_solvers_and_layout_switches(O, spk):
    A = spk.powerlaw_spd(60000, 3)
    b = spk.uniform_pm1(9, A.nrows)
    its = []
    for code in (1, -(16 * 65536 + 8192), 8):
        A.set_layout(code)
        assert A.panels == code
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.conjugate_gradient(b, x, spk.CGConfig(1e-10, 1000, 1000), A,
                                   spk.jacobi(A))
        assert s.converged and spk.true_relres(A, b, x) < 1e-9
        its.append(s.computed_iterations)
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.bicgstab(b, x, spk.BicgstabConfig(1e-10, 1000), A, spk.jacobi(A))
        assert s.converged
    assert max(its) - min(its) <= 1
    # a row whose columns do not ascend is rejected (the store is unchanged)
    U = spk.CsrMatrix.from_host(2, 3, [0, 2, 3], [2, 0, 1], [1.0, 2.0, 3.0])
    with pytest.raises(spk.InvalidArgument, match="ascending"):
        U.set_window_tiles(8, 256)
    assert U.panels == 1


def test_window_tiles_edge_cases(O, spk):
    """sparse_blas_test.cpp:56-95's empty, rectangular and identity cases
    under window tiles (an empty store stays plain), and a store whose
    every row is long: the tiles hold no entry and the long-row store sums
    every row."""
    E = spk.CsrMatrix.from_oracle(O.make_matrix(3, 3, [], [], []))
    E.set_window_tiles(8, 256)
    assert E.panels == 1
    x = spk.DeviceVector.from_numpy(np.array([1., 2., 3.]))
    assert spk.multiply(E, x).numpy().tolist() == [0., 0., 0.]
    R = spk.CsrMatrix.from_oracle(O.make_matrix(2, 3, [0, 0, 1], [0, 2, 1],
                                                [1., 2., 3.]))
    R.set_window_tiles(8, 256)
    assert spk.multiply(R, x).numpy().tolist() == [7., 6.]
    I = spk.CsrMatrix.from_oracle(O.make_matrix(3, 3, [0, 1, 2], [0, 1, 2],
                                                [1., 1., 1.]))
    I.set_window_tiles(16, 8192)
    assert spk.multiply(I, spk.DeviceVector.from_numpy(
        np.array([4., 5., 6.]))).numpy().tolist() == [4., 5., 6.]
    rng = np.random.default_rng(11)
    n, nc = 40, 3000  # one window of 8192 columns: rows over 256 are long
    cols = [np.sort(rng.choice(nc, size=400, replace=False)) for _ in range(n)]
    rp = np.arange(n + 1) * 400
    ci = np.concatenate(cols).astype(np.int64)
    v = rng.uniform(-1, 1, len(ci))
    L = spk.CsrMatrix.from_host(n, nc, rp, ci, v)
    L.set_window_tiles(16, 8192)
    xl = O.uniform_pm1(2, nc)
    y = spk.multiply(L, spk.DeviceVector.from_numpy(xl)).numpy()
    ref = O.spmv(O.Csr(n, nc, rp, ci, v), xl)
    absrow = np.abs(O.Csr(n, nc, rp, ci, v).to_scipy()) @ np.abs(xl)
    assert np.all(np.abs(y - ref) <= 1e-12 * absrow)
    yd = spk.DeviceVector.zeros(n)
    r = spk.DeviceVector.zeros(1)
    with pytest.raises(spk.InvalidArgument):
        L.spmv_dot(spk.DeviceVector.from_numpy(xl), yd, r)  # not square
