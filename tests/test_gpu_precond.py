"""GPU parity of the preconditioners: jacobi / jacobi_apply (jacobi.hpp),
ilu0 / milu0 / ilu_apply (ilu.hpp) and the triangular sweeps
(triangular_solve.hpp) against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dv(spk, a):
    return spk.DeviceVector.from_numpy(np.asarray(a))


def test_jacobi_known_values(O, spk):
    # jacobi_test.cpp:50-95
    A = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 4., 5., 10.])
    J = spk.jacobi(spk.CsrMatrix.from_oracle(A))
    z = spk.DeviceVector.zeros(4)
    spk.jacobi_apply(J, dv(spk, [6., 12., 20., 30.]), z)
    assert z.numpy().tolist() == [3.0, 3.0, 4.0, 3.0]


def test_jacobi_bit_exact(O, spk):
    A = O.convdiff_3d(11, 9, 7, 1.0, 0.37, 0.11, 0.05)
    J = spk.jacobi(spk.CsrMatrix.from_oracle(A))
    r = O.uniform_pm1(9, A.nrows)
    z = spk.DeviceVector.zeros(A.nrows)
    J.apply(dv(spk, r), z)
    assert np.array_equal(z.numpy(), O.jacobi_apply(O.jacobi(A), r))


def test_jacobi_zero_diagonal_throws(O, spk):
    # jacobi_test.cpp:273-282
    A = O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 0, 2], [2., 1., 1., 3.])
    with pytest.raises(spk.InvalidArgument):
        spk.jacobi(spk.CsrMatrix.from_oracle(A))


def _factor_cases(O):
    return [
        O.tridiagonal_matrix(9, -1.0, 4.0, -1.0),
        O.nonsymmetric_sample(),
        O.convdiff_centered_2d(12, 9, 1.0, 0.3, 0.2, 5.0),
        O.convection_diffusion_2d(17, 13, 1.0, 0.45, 0.2),
        O.convdiff_3d(7, 6, 5, 1.0, 0.4, 0.2, 0.1),
        O.poisson_2d(10, 10, 0.5),
    ]


@pytest.mark.parametrize("modified", [False, True])
def test_ilu0_factors_bit_exact(O, spk, modified):
    """The level-scheduled device factorization performs the reference's
    IKJ updates in the same order with no fused multiply-add: L and U match
    bit for bit (patterns and values)."""
    for A in _factor_cases(O):
        Lr, Ur = O.ilu0(A, modified)
        F = (spk.milu0 if modified else spk.ilu0)(spk.CsrMatrix.from_oracle(A))
        (lrp, lci, lv), (urp, uci, uv) = F.factors()
        assert np.array_equal(lrp, Lr.row_ptr) and np.array_equal(lci, Lr.col_ind)
        assert np.array_equal(urp, Ur.row_ptr) and np.array_equal(uci, Ur.col_ind)
        assert np.array_equal(lv, Lr.vals)
        assert np.array_equal(uv, Ur.vals)


def test_ilu0_apply_bit_exact(O, spk):
    for A in _factor_cases(O):
        Lr, Ur = O.ilu0(A)
        F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
        r = O.uniform_pm1(11, A.nrows)
        z = spk.DeviceVector.zeros(A.nrows)
        spk.ilu_apply(F, dv(spk, r), z)
        z_ref = O.backward_solve(Ur, O.forward_solve(Lr, r))
        assert np.array_equal(z.numpy(), z_ref)


def test_ilu0_tridiagonal_exact_solve(O, spk):
    # ilu_test.cpp:93-107: no fill, so ilu_apply is an exact solve
    A = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    b = O.spmv(A, np.array([1., 2., 3., 4.]))
    F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
    z = spk.DeviceVector.zeros(4)
    spk.ilu_apply(F, dv(spk, b), z)
    assert np.allclose(z.numpy(), [1., 2., 3., 4.], atol=1e-12, rtol=0)


def test_ilu0_errors(O, spk):
    # ilu_test.cpp:274-282 (non-square); ilu.hpp:63-65 (missing diagonal)
    R = O.make_matrix(3, 4, [0, 1, 2], [0, 1, 2], [1., 2., 3.])
    with pytest.raises(spk.InvalidArgument):
        spk.ilu0(spk.CsrMatrix.from_oracle(R))
    M = O.make_matrix(3, 3, [0, 1, 1, 2], [0, 0, 2, 2], [1., 1., 1., 1.])
    with pytest.raises(spk.InvalidArgument):
        spk.ilu0(spk.CsrMatrix.from_oracle(M))


def test_ilu0_fp32_apply(O, spk):
    A = O.convdiff_3d(9, 8, 7, 1.0, 0.5, 0.25, 0.125)
    Lr, Ur = O.ilu0(A)
    F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
    r = O.uniform_pm1(12, A.nrows)
    z = spk.DeviceVector.zeros(A.nrows, np.float32)
    F.apply(dv(spk, r.astype(np.float32)), z)
    z_ref = O.backward_solve(Ur, O.forward_solve(Lr, r))
    assert np.linalg.norm(z.numpy() - z_ref) <= 1e-5 * np.linalg.norm(z_ref)


def test_ilu0_levels(O, spk):
    # natural-order 7-point stencil: i+j+k hyperplanes
    F = spk.ilu0(spk.CsrMatrix.from_oracle(O.convdiff_3d(6, 5, 4, 1., .1, .1, .1)))
    assert F.levels() == (6 + 5 + 4 - 2, 6 + 5 + 4 - 2)


@pytest.mark.parametrize("schedule", ["levels", "sync-free"])
def test_ilu0_schedules_bit_exact(O, spk, schedule):
    """Both sweep schedules run the reference's arithmetic: applications,
    standalone sweeps, repeated and in-place use are bit-identical to the
    oracle."""
    for A in (O.convdiff_3d(13, 11, 9, 1.0, 0.3, 0.2, 0.1),
              O.convdiff_centered_2d(17, 15, 1.0, 0.3, 0.2, 3.0)):
        Lr, Ur = O.ilu0(A)
        F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
        F.schedule = schedule
        assert F.schedule == schedule
        for seed in (1, 2):  # repeated applications
            r = O.uniform_pm1(seed, A.nrows)
            z = spk.DeviceVector.zeros(A.nrows)
            spk.ilu_apply(F, dv(spk, r), z)
            z_ref = O.backward_solve(Ur, O.forward_solve(Lr, r))
            assert np.array_equal(z.numpy(), z_ref)
        zz = dv(spk, r)              # in place
        F.apply(zz, zz)
        assert np.array_equal(zz.numpy(), z_ref)
        y = spk.DeviceVector.zeros(A.nrows)
        F.forward(dv(spk, r), y)
        assert np.array_equal(y.numpy(), O.forward_solve(Lr, r))
        w = spk.DeviceVector.zeros(A.nrows)
        F.backward(y, w)
        assert np.array_equal(w.numpy(), z_ref)
        # the standalone sweeps write a separate output: an overlapping
        # call is rejected before anything is launched, on both schedules
        for sweep in (F.forward, F.backward):
            with pytest.raises(spk.InvalidArgument, match="overlap"):
                sweep(y, y)
        assert np.array_equal(y.numpy(), O.forward_solve(Lr, r))
        # switching schedules on a used preconditioner keeps it exact
        F.schedule = "sync-free" if schedule == "levels" else "levels"
        z2 = spk.DeviceVector.zeros(A.nrows)
        F.apply(dv(spk, r), z2)
        assert np.array_equal(z2.numpy(), z_ref)
        F.schedule = "auto"


def test_ilu0_auto_schedule_picks_levels_for_wide_levels(O, spk):
    A = spk.convdiff_3d(48, 48, 48, 1.0, 0.1, 0.05, 0.02)
    F = spk.ilu0(A)
    assert F.schedule == "levels"          # 110592 rows, 142 levels
    S = spk.ilu0(spk.CsrMatrix.from_oracle(O.tridiagonal_matrix(300, -1., 4., -1.)))
    assert S.schedule == "sync-free"


@pytest.mark.parametrize("k", [1, 2, 3, 4, 7])
def test_ilu0_jacobi_sweeps_match_oracle(O, spk, k):
    """Approximate apply (k Jacobi sweeps per triangle; no reference
    counterpart): on stencil rows the device sums each row in storage order
    and divides like the oracle, so FP64 results are bit-identical; FP32
    within FP32 rounding; in-place use and switching back to exact solves."""
    for A in (O.convdiff_3d(13, 11, 9, 1.0, 0.3, 0.2, 0.1),
              O.convdiff_centered_2d(17, 15, 1.0, 0.3, 0.2, 3.0)):
        Lr, Ur = O.ilu0(A)
        F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
        F.jacobi_sweeps = k
        assert F.jacobi_sweeps == k
        for seed in (1, 2):
            r = O.uniform_pm1(seed, A.nrows)
            z = spk.DeviceVector.zeros(A.nrows)
            F.apply(dv(spk, r), z)
            z_ref = O.ilu_jacobi_apply(Lr, Ur, k, r)
            assert np.array_equal(z.numpy(), z_ref)
        zz = dv(spk, r)
        F.apply(zz, zz)
        assert np.array_equal(zz.numpy(), z_ref)
        z32 = spk.DeviceVector.zeros(A.nrows, np.float32)
        F.apply(dv(spk, r.astype(np.float32)), z32)
        ref32 = O.ilu_jacobi_apply(Lr, Ur, k, r.astype(np.float32))
        assert np.allclose(z32.numpy(), ref32, rtol=0,
                           atol=1e-5 * np.abs(ref32).max())
        # standalone sweeps stay exact; sweeps = 0 restores ilu_apply
        y = spk.DeviceVector.zeros(A.nrows)
        F.forward(dv(spk, r), y)
        assert np.array_equal(y.numpy(), O.forward_solve(Lr, r))
        F.jacobi_sweeps = 0
        z0 = spk.DeviceVector.zeros(A.nrows)
        F.apply(dv(spk, r), z0)
        assert np.array_equal(z0.numpy(),
                              O.backward_solve(Ur, O.forward_solve(Lr, r)))


def test_ilu0_jacobi_sweeps_reach_exact_at_depth(O, spk):
    """The strict triangles are nilpotent: with as many sweeps as dependency
    levels the approximate apply is the exact ilu_apply."""
    A = O.convdiff_3d(7, 6, 5, 1.0, 0.3, 0.2, 0.1)
    Lr, Ur = O.ilu0(A)
    F = spk.ilu0(spk.CsrMatrix.from_oracle(A))
    F.jacobi_sweeps = max(F.levels())
    r = O.uniform_pm1(3, A.nrows)
    z = spk.DeviceVector.zeros(A.nrows)
    F.apply(dv(spk, r), z)
    ex = O.backward_solve(Ur, O.forward_solve(Lr, r))
    assert np.allclose(z.numpy(), ex, rtol=0, atol=1e-13 * np.abs(ex).max())
    with pytest.raises(spk.InvalidArgument):
        F.jacobi_sweeps = -1
    with pytest.raises(spk.InvalidArgument):  # not an ILU preconditioner
        spk.Ilu0.jacobi_sweeps.fset(spk.jacobi(spk.CsrMatrix.from_oracle(A)), 2)
