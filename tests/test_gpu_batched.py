"""Batched persistent Jacobi-PCG on the GPU: the device-built batch is
bit-identical to the oracle's systems; every system reaches the FP64 true
residual tolerance (checked independently by the oracle's SpMV); the FP64
mode's per-system iteration counts match the oracle's Jacobi-PCG within one;
graph replays and the autotuner's launch shapes give identical results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED, SMIN, SMAX, COUNT = 77, 6, 40, 48


@pytest.fixture(scope="module")
def batch(spk):
    return spk.BatchedSystems.diffusion(COUNT, SEED, SMIN, SMAX)


def sys_csr(O, rp, ci, v, r0, r1):
    a, b = rp[r0], rp[r1]
    return O.Csr(r1 - r0, r1 - r0, rp[r0:r1 + 1] - a, ci[a:b] - r0,
                 v[a:b].copy())


def test_generator_bit_exact(O, spk, batch):
    ro = batch.row_offsets()
    rp, ci, v = batch.matrix.download()
    assert batch.count == COUNT and ro[-1] == batch.total_rows
    for k in range(COUNT):
        R = O.batched_diffusion(SEED, k, SMIN, SMAX)
        S = sys_csr(O, rp, ci, v, ro[k], ro[k + 1])
        assert np.array_equal(S.row_ptr, R.row_ptr), k
        assert np.array_equal(S.col_ind, R.col_ind), k
        assert np.array_equal(S.vals.view(np.uint64), R.vals.view(np.uint64)), k


def _rhs(spk, batch):
    return spk.uniform_pm1(5, batch.total_rows)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_every_system_converges(O, spk, batch, prec):
    b = _rhs(spk, batch)
    x = spk.DeviceVector.zeros(batch.total_rows)
    if prec == "f32":
        cfg = spk.BatchConfig(1e-8, 1e-4, 30, 5000, spk.F32)
    else:
        cfg = spk.BatchConfig(1e-8, 1e-8, 1, 5000, spk.F64)
    s = batch.solve(cfg, b, x)
    assert s["converged"] == COUNT and s["max_relres"] < 1e-8
    it, ou, rr, cv = batch.results()
    assert np.all(cv == 1) and np.all(rr < 1e-8)
    ro = batch.row_offsets()
    rp, ci, v = batch.matrix.download()
    bh, xh = b.numpy(), x.numpy()
    for k in range(0, COUNT, 5):
        S = sys_csr(O, rp, ci, v, ro[k], ro[k + 1])
        bk, xk = bh[ro[k]:ro[k + 1]], xh[ro[k]:ro[k + 1]]
        rel = O.true_relres(S, bk, xk)
        assert rel < 1e-8
        assert abs(rel - rr[k]) <= 1e-3 * rel + 1e-15
        if prec == "f64":
            xr, sr = O.cg(S, bk, tol=1e-8, restart_iterations=5000,
                          max_iterations=5000, left="jacobi")
            # the device judges on the true residual, the reference on the
            # recursive one: equal up to a step either way
            assert abs(it[k] - sr.computed_iterations) <= 2, k
            assert np.linalg.norm(xk - xr) <= 1e-6 * np.linalg.norm(xr)


def test_graph_replay_and_tunings_agree(spk, batch):
    b = _rhs(spk, batch)
    cfg = spk.BatchConfig(1e-8, 1e-4, 30, 5000, spk.F32)
    x0 = spk.DeviceVector.zeros(batch.total_rows)
    batch.tuning = (256, 2, 0)
    batch.solve(cfg, b, x0)
    ref = x0.numpy()
    it0 = batch.results()[0].copy()
    x1 = spk.DeviceVector.zeros(batch.total_rows)
    for rep in range(3):  # capture, then replays
        batch.solve(cfg, b, x1, use_graph=True)
        assert np.array_equal(x1.numpy(), ref), rep
        assert np.array_equal(batch.results()[0], it0)
    for t in [(512, 1, 4096), (1024, 2, 0), (256, 2, 16384)]:
        batch.tuning = t
        x2 = spk.DeviceVector.zeros(batch.total_rows)
        batch.solve(cfg, b, x2)
        # same arithmetic per system except the reduction tree width
        assert np.all(batch.results()[3] == 1)
        assert np.linalg.norm(x2.numpy() - ref) <= 1e-6 * np.linalg.norm(ref)


def test_autotune(spk, batch):
    b = _rhs(spk, batch)
    x = spk.DeviceVector.zeros(batch.total_rows)
    cfg = spk.BatchConfig(1e-8, 1e-4, 30, 5000, spk.F32)
    chosen, tried = batch.autotune(cfg, b, x, sample=16)
    assert len(tried) >= 10 and chosen in [t for t, _ in tried]
    assert min(ms for _, ms in tried) == dict(tried)[chosen]
    assert batch.tuning == chosen
    s = batch.solve(cfg, b, x)
    assert s["converged"] == COUNT


def test_tuning_record_round_trip(spk, batch, tmp_path):
    """The autotuner's choice as a small JSON record keyed by the GPU
    target, applied to another batch; a record from another target or a
    foreign file is rejected."""
    import json
    b = _rhs(spk, batch)
    x = spk.DeviceVector.zeros(batch.total_rows)
    cfg = spk.BatchConfig(1e-8, 1e-4, 30, 5000, spk.F32)
    chosen, tried = batch.autotune(cfg, b, x, sample=8)
    path = tmp_path / "tuning.json"
    batch.save_tuning(str(path), tried)
    doc = json.loads(path.read_text())
    assert doc["gfx_target"].startswith("gfx") and len(doc["tried"]) == len(tried)
    other = spk.BatchedSystems.diffusion(6, SEED + 1, SMIN, SMAX)
    other.tuning = (256, 1, 0)
    assert other.load_tuning(str(path)) == chosen == other.tuning
    doc["gfx_target"] = "gfx000"
    path.write_text(json.dumps(doc))
    with pytest.raises(spk.InvalidArgument, match="tuned on gfx000"):
        other.load_tuning(str(path))
    path.write_text("{}")
    with pytest.raises(spk.InvalidArgument, match="not a tuning record"):
        other.load_tuning(str(path))


def test_from_blocks_and_errors(O, spk):
    A1 = O.batched_diffusion(3, 0, 4, 4)
    A2 = O.tridiagonal_matrix(5, -1.0, 3.0, -1.0)
    n1, n2 = A1.nrows, A2.nrows
    rp = np.concatenate([A1.row_ptr, A2.row_ptr[1:] + A1.nnz])
    ci = np.concatenate([A1.col_ind, A2.col_ind + n1])
    v = np.concatenate([A1.vals, A2.vals])
    A = spk.CsrMatrix.from_host(n1 + n2, n1 + n2, rp, ci, v)
    B = spk.BatchedSystems.from_blocks(A, [0, n1, n1 + n2])
    b = spk.DeviceVector.from_numpy(np.ones(n1 + n2))
    x = spk.DeviceVector.zeros(n1 + n2)
    s = B.solve(spk.BatchConfig(1e-10, 1e-10, 1, 100, spk.F64), b, x)
    assert s["converged"] == 2
    with pytest.raises(spk.InvalidArgument):   # block 0 would couple to 1
        spk.BatchedSystems.from_blocks(A, [0, n1 - 1, n1 + n2])
    with pytest.raises(spk.InvalidArgument):
        B.tuning = (300, 1, 0)


def _arrow(O, n):
    """Arrowhead SPD matrix: row 0 holds all n entries (wider than any
    sliced-ELL width), the rest tridiagonal plus the first column."""
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n):
            if i == j:
                v = 3.0 * n
            elif i == 0 or j == 0 or abs(i - j) == 1:
                v = -1.0
            else:
                continue
            rows.append(i), cols.append(j), vals.append(v)
    return O.make_matrix(n, n, rows, cols, vals)


@pytest.mark.parametrize("kind", ["w3", "w5", "w7", "w9", "csr"])
def test_every_row_width(O, spk, kind):
    """Batches whose widest row takes each compiled sliced-ELL width (rows
    padded to it), and one with a 12-entry row that keeps the CSR loop:
    per-system FP64 iteration counts match the oracle's Jacobi-PCG, FP32
    reaches the FP64 true-residual tolerance."""
    mk = {"w3": lambda: [O.tridiagonal_matrix(40, -1.0, 2.5, -1.0),
                         O.tridiagonal_matrix(17, -1.0, 3.0, -1.0)],
          "w5": lambda: [O.poisson_2d(9, 7, 0.1), O.tridiagonal_matrix(30, -1.0, 2.2, -1.0)],
          "w7": lambda: [O.poisson_3d(6, 5, 4), O.poisson_2d(8, 8, 0.05)],
          "w9": lambda: [O.aniso_q1_2d(12, 5, 1.5), O.poisson_3d(5, 5, 5)],
          "csr": lambda: [_arrow(O, 12), O.poisson_2d(6, 6, 0.1)]}[kind]
    blocks = mk()
    offs, rps, cis, vs, nz = [0], [np.zeros(1, np.int32)], [], [], 0
    for S in blocks:
        r0 = offs[-1]
        rps.append((S.row_ptr[1:] + nz).astype(np.int32))
        cis.append((S.col_ind + r0).astype(np.int32))
        vs.append(S.vals)
        offs.append(r0 + S.nrows)
        nz += S.nnz
    n = offs[-1]
    A = spk.CsrMatrix.from_host(n, n, np.concatenate(rps), np.concatenate(cis),
                                np.concatenate(vs))
    B = spk.BatchedSystems.from_blocks(A, offs)
    bh = O.uniform_pm1(9, n)
    for prec in ("f64", "f32"):
        x = spk.DeviceVector.zeros(n)
        if prec == "f64":
            cfg = spk.BatchConfig(1e-9, 1e-9, 1, 3000, spk.F64)
        else:
            cfg = spk.BatchConfig(1e-9, 1e-3, 40, 3000, spk.F32)
        s = B.solve(cfg, spk.DeviceVector.from_numpy(bh), x)
        assert s["converged"] == len(blocks), (kind, prec)
        it = B.results()[0]
        xh = x.numpy()
        for k, S in enumerate(blocks):
            bk, xk = bh[offs[k]:offs[k + 1]], xh[offs[k]:offs[k + 1]]
            assert O.true_relres(S, bk, xk) < 1e-9, (kind, prec, k)
            if prec == "f64":
                _, sr = O.cg(S, bk, tol=1e-9, restart_iterations=3000,
                             max_iterations=3000, left="jacobi")
                assert abs(it[k] - sr.computed_iterations) <= 2, (kind, k)


def test_wide_band_system_keeps_32_bit_columns(O, spk):
    """The sliced-ELL copy stores each column as a 16-bit offset from its
    row when every column lies within 32767 of its row (the diffusion
    batches above); a system coupling rows 40,000 apart keeps the 32-bit
    columns.  Both solve to the oracle's Jacobi-PCG: same counts within the
    true-residual step, same x."""
    n1 = 40001
    rows, cols, vals = [], [], []
    for i in range(n1):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n1:
                rows.append(i); cols.append(j); vals.append(3.0 if i == j else -1.0)
    rows += [0, n1 - 1]; cols += [n1 - 1, 0]; vals += [-0.5, -0.5]
    W = O.make_matrix(n1, n1, rows, cols, vals)
    D = O.batched_diffusion(11, 0, 20, 20)
    for blocks in ((W, D), (D,)):
        n = sum(B.nrows for B in blocks)
        rp, ci, v, off = [np.zeros(1, np.int64)], [], [], [0]
        base = 0
        for B in blocks:
            rp.append(B.row_ptr[1:] + rp[-1][-1])
            ci.append(B.col_ind + off[-1])
            v.append(B.vals)
            off.append(off[-1] + B.nrows)
        A = spk.CsrMatrix.from_host(n, n, np.concatenate(rp), np.concatenate(ci),
                                    np.concatenate(v))
        batch = spk.BatchedSystems.from_blocks(A, off)
        b = spk.uniform_pm1(9, n)
        x = spk.DeviceVector.zeros(n)
        s = batch.solve(spk.BatchConfig(1e-8, 1e-8, 1, 20000, spk.F64), b, x)
        assert s["converged"] == len(blocks)
        it = batch.results()[0]
        bh, xh = b.numpy(), x.numpy()
        for k, B in enumerate(blocks):
            bk, xk = bh[off[k]:off[k + 1]], xh[off[k]:off[k + 1]]
            xr, sr = O.cg(B, bk, tol=1e-8, restart_iterations=20000,
                          max_iterations=20000, left="jacobi")
            assert abs(it[k] - sr.computed_iterations) <= 2, k
            assert np.linalg.norm(xk - xr) <= 1e-6 * np.linalg.norm(xr)
