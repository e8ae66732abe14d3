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
