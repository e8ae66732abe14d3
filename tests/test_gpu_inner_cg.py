"""The inner-PCG preconditioner (spk_precond_create_inner_cg) in FGMRES's
flexible slot (fgmres.hpp:246-256): one application is exactly a
conjugate_gradient solve (conjugate_gradient.hpp:122-312) from zero with the
inner preconditioner, and FGMRES(30) with it reaches the FP64 true-residual
target."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _learned_sym(spk, A):
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(16)]
    model = spk.SpaiModel(1)
    model.train(mats, epochs=4, learning_rate=2e-3, shuffle_seed=3)
    return spk.spai(model.predict(A).symmetric_part())


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_one_application_is_a_cg_solve(spk, prec):
    A = spk.aniso_q1_2d(96, 11, 1.5)
    M = _learned_sym(spk, A)
    p = spk.F32 if prec == "f32" else spk.F64
    dt = np.float32 if prec == "f32" else np.float64
    P = spk.inner_cg(A, M, iterations=37, tolerance=1e-3, precision=p)
    r = spk.uniform_pm1(5, A.nrows)
    z = spk.DeviceVector.zeros(A.nrows)
    P.apply(r, z)
    # the same solve through conjugate_gradient, in the inner precision
    rr = r.astype(dt)
    zz = spk.DeviceVector.zeros(A.nrows, dt)
    s = spk.conjugate_gradient(rr, zz, spk.CGConfig(1e-3, 37, 37, False, p),
                               A, M)
    assert P.inner_iterations == s.computed_iterations <= 37
    assert np.array_equal(z.numpy(), zz.numpy().astype(np.float64))
    # in place: r is read before z is written
    P.apply(r, r)
    assert np.array_equal(r.numpy(), z.numpy())
    assert P.inner_iterations == 2 * s.computed_iterations


def test_fgmres_with_inner_pcg_reaches_target(spk):
    A = spk.aniso_q1_2d(128, 4242, 1.5)
    b = spk.uniform_pm1(42, A.nrows)
    M = _learned_sym(spk, A)
    P = spk.inner_cg(A, M, iterations=60, precision=spk.F32)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.FgmresConfig(1e-8, 30, 500, False, spk.F64, spk.ORTHO_CGS2)
    s = spk.fgmres(b, x, cfg, A, P)
    assert s.converged and spk.true_relres(A, b, x) < 1e-8
    # each outer iteration ran one inner solve of at most 60 steps
    assert 0 < P.inner_iterations <= 60 * s.computed_iterations
    x2 = spk.DeviceVector.zeros(A.nrows)
    s2 = spk.fgmres(b, x2, cfg, A, M)  # the same M applied once per step
    assert s2.computed_iterations > s.computed_iterations


def test_inner_pcg_rejected_where_a_fixed_operator_is_needed(spk):
    A = spk.poisson_2d(20, 20, 0.5)
    P = spk.inner_cg(A, spk.jacobi(A), iterations=5)
    b = spk.uniform_pm1(1, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    with pytest.raises(spk.InvalidArgument, match="inner-PCG"):
        spk.conjugate_gradient(b, x, spk.CGConfig(1e-8, 100, 100), A, P)
    with pytest.raises(spk.InvalidArgument):
        spk.inner_cg(A, P, iterations=5)  # no nesting
    with pytest.raises(spk.InvalidArgument):
        spk.inner_cg(A, None, iterations=0)


def _ir_cfg(spk, tol=1e-6):
    return spk.IrConfig(tolerance=tol, max_outer=30, inner_tolerance=0.1,
                        restart=0, max_inner=20000,
                        inner_method=spk.IR_CG_RELIABLE)


def test_inner_ir_application_is_an_ir_solve(spk):
    """spk_precond_create_inner_ir: one application is spk_ir_solve from
    zero (FP32 PCG with FP64 reliable updates, the learned inverse's
    symmetric part inside), bit for bit; in place; and for an FP32 caller
    the FP64 result rounded."""
    A = spk.aniso_q1_2d(96, 11, 1.5)
    M = _learned_sym(spk, A)
    cfg = _ir_cfg(spk)
    P = spk.inner_ir(A, M, cfg)
    r = spk.uniform_pm1(5, A.nrows)
    z = spk.DeviceVector.zeros(A.nrows)
    P.apply(r, z)
    zz = spk.DeviceVector.zeros(A.nrows)
    s = spk.ir_solve(r, zz, cfg, A, M)
    assert s.converged and s.relres < 1e-6
    assert P.inner_iterations == s.inner_iterations
    assert np.array_equal(z.numpy(), zz.numpy())
    P.apply(r, r)  # r is read before z is written
    assert np.array_equal(r.numpy(), z.numpy())
    r2 = spk.uniform_pm1(5, A.nrows)
    r32 = r2.astype(np.float32)
    z32 = spk.DeviceVector.zeros(A.nrows, np.float32)
    P.apply(r32, z32)
    b64 = spk.DeviceVector.from_numpy(r32.numpy().astype(np.float64))
    z64 = spk.DeviceVector.zeros(A.nrows)
    spk.ir_solve(b64, z64, cfg, A, M)
    assert np.array_equal(z32.numpy(), z64.numpy().astype(np.float32))


def test_fgmres_with_inner_ir_reaches_target(spk):
    A = spk.aniso_q1_2d(128, 4242, 1.5)
    b = spk.uniform_pm1(42, A.nrows)
    M = _learned_sym(spk, A)
    P = spk.inner_ir(A, M, _ir_cfg(spk, 1e-4))
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.FgmresConfig(1e-8, 30, 500, False, spk.F64, spk.ORTHO_CGS2)
    s = spk.fgmres(b, x, cfg, A, P)
    assert s.converged and spk.true_relres(A, b, x) < 1e-8
    # each application cuts its residual 1e4-fold: a handful of outer steps
    assert 1 <= s.computed_iterations <= 4
    assert P.inner_iterations > 0


def test_inner_ir_rejections(spk):
    A = spk.poisson_2d(20, 20, 0.5)
    J = spk.jacobi(A)
    P = spk.inner_ir(A, J, _ir_cfg(spk))
    b = spk.uniform_pm1(1, A.nrows)
    x = spk.DeviceVector.zeros(A.nrows)
    with pytest.raises(spk.InvalidArgument, match="inner-solve"):
        spk.conjugate_gradient(b, x, spk.CGConfig(1e-8, 100, 100), A, P)
    with pytest.raises(spk.InvalidArgument, match="inner-solve"):
        spk.ir_solve(b, x, _ir_cfg(spk), A, P)
    bad = _ir_cfg(spk)
    bad.inner_method = spk.IR_GMRES  # would share the outer GMRES's vectors
    with pytest.raises(spk.InvalidArgument, match="inner_method"):
        spk.inner_ir(A, J, bad)
    with pytest.raises(spk.InvalidArgument, match="nest"):
        spk.inner_ir(A, P, _ir_cfg(spk))
    with pytest.raises(spk.InvalidArgument, match="nest"):
        spk.inner_cg(A, P, iterations=5)
