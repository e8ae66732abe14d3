"""The learned preconditioner's guarded solve (spk_learned_solve): FGMRES
(fgmres.hpp:452-470) with the network's M, and the fallback -- Jacobi-PCG
(conjugate_gradient.hpp:122-312, jacobi.hpp:25-50) for symmetric A,
ILU(0)-GMRES (gmres.hpp:99-375, ilu.hpp:35-179) otherwise -- when M is
poisoned, when FGMRES stalls or misses the tolerance.  The FP64 true
residual is the judge.  Poisoned networks live only in M's values: every
kernel runs to completion on finite indices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# canonical parameter layout: W1[24][64] b1[64] W2[64][64] b2[64] W3[64][9] b3[9]
W3 = 24 * 64 + 64 + 64 * 64 + 64
B3 = W3 + 64 * 9


def _aniso(spk, side=64, seed=1000):
    A = spk.aniso_q1_2d(side, seed, 1.5)
    return A, spk.uniform_pm1(7, A.nrows)


def _trained(spk):
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(16)]
    model = spk.SpaiModel(1)
    model.train(mats, epochs=4, learning_rate=2e-3, shuffle_seed=3)
    return model


def test_learned_path_converges_without_fallback(spk):
    A, b = _aniso(spk)
    M = _trained(spk).preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.learned_solve(b, x, spk.LearnedConfig(tolerance=1e-8), A, M)
    assert s.method == "learned_fgmres" and not s.fell_back
    assert s.reasons == [] and s.converged
    assert s.learned_iterations > 0 and s.fallback_iterations == 0
    assert s.relres == s.learned_relres < 1e-8
    assert abs(spk.true_relres(A, b, x) - s.relres) <= 1e-3 * s.relres
    # the same FGMRES run outside the guard
    x2 = spk.DeviceVector.zeros(A.nrows)
    s2 = spk.fgmres(b, x2, spk.FgmresConfig(1e-8, 30, 50000, False, spk.F64,
                                            spk.ORTHO_CGS2), A, M)
    assert s2.computed_iterations == s.learned_iterations
    assert np.array_equal(x.numpy(), x2.numpy())


def test_learned_path_with_single_pass_gram_schmidt(spk):
    """The guarded solve with classical Gram-Schmidt once (SPK_ORTHO_CGS1):
    its FGMRES is the same as outside the guard, and the FP64 true residual
    still decides (the estimate may drift with a less orthogonal basis)."""
    A, b = _aniso(spk)
    M = _trained(spk).preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.LearnedConfig(tolerance=1e-8, orthogonalization=spk.ORTHO_CGS1)
    s = spk.learned_solve(b, x, cfg, A, M)
    assert s.converged and spk.true_relres(A, b, x) < 1e-8
    assert abs(spk.true_relres(A, b, x) - s.relres) <= 1e-3 * s.relres
    if not s.fell_back:
        x2 = spk.DeviceVector.zeros(A.nrows)
        s2 = spk.fgmres(b, x2, spk.FgmresConfig(1e-8, 30, 50000, False, spk.F64,
                                                spk.ORTHO_CGS1), A, M)
        assert s2.computed_iterations == s.learned_iterations
        assert np.array_equal(x.numpy(), x2.numpy())


def test_nan_output_layer_falls_back_to_jacobi_pcg(spk):
    A, b = _aniso(spk)
    model = _trained(spk)
    p = model.params
    p[W3:B3] = np.nan
    model.params = p
    M = model.preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.learned_solve(b, x, spk.LearnedConfig(tolerance=1e-8), A, M)
    assert s.fell_back and s.method == "jacobi_pcg"
    # every predicted entry is NaN, the diagonal included
    assert s.reasons == ["bad_values", "bad_diagonal"]
    assert s.bad_values == A.nnz and s.bad_diagonal == A.nrows
    assert s.learned_iterations == 0 and np.isnan(s.learned_relres)
    assert s.converged and s.relres < 1e-8
    assert spk.true_relres(A, b, x) < 1e-8
    # the fallback is Jacobi-PCG from x0: the count of the plain solve
    x2 = spk.DeviceVector.zeros(A.nrows)
    s2 = spk.conjugate_gradient(b, x2, spk.CGConfig(1e-8, 50000, 50000), A,
                                spk.jacobi(A))
    assert s.fallback_iterations == s2.computed_iterations


def test_zero_diagonal_output_falls_back(spk):
    """A zero output layer except b3[diagonal slot] = -1: every predicted
    diagonal entry is (1 + y) / a_ii = 0."""
    A, b = _aniso(spk)
    model = spk.SpaiModel(1)
    p = model.params
    p[W3:] = 0.0
    p[B3 + 4] = -1.0
    model.params = p
    M = model.preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.learned_solve(b, x, spk.LearnedConfig(tolerance=1e-8), A, M)
    assert s.fell_back and "bad_diagonal" in s.reasons
    assert s.bad_diagonal == A.nrows and s.bad_values == 0
    assert s.converged and spk.true_relres(A, b, x) < 1e-8


def test_zero_output_layer_stalls_then_falls_back(spk):
    """A zero output layer predicts the Jacobi inverse, which FGMRES(30)
    turns into a slow solve on this matrix: with a strict stall rule (each
    cycle must cut the residual estimate 100-fold) it is stopped after its
    third cycle and Jacobi-PCG finishes from the learned x."""
    A, b = _aniso(spk, side=128)
    model = spk.SpaiModel(1)
    p = model.params
    p[W3:] = 0.0
    model.params = p
    M = model.preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    cfg = spk.LearnedConfig(tolerance=1e-8, stall_factor=0.01, stall_cycles=2)
    s = spk.learned_solve(b, x, cfg, A, M)
    assert s.fell_back and s.method == "jacobi_pcg"
    assert s.reasons == ["stalled"]
    assert s.learned_iterations == 90
    assert s.learned_relres > 1e-8
    assert s.converged and spk.true_relres(A, b, x) < 1e-8


def test_budget_spent_falls_back(spk):
    A, b = _aniso(spk, side=128)
    M = spk.SpaiModel(1).preconditioner(A)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.learned_solve(b, x, spk.LearnedConfig(tolerance=1e-8,
                                                  max_iterations=45), A, M)
    assert s.reasons == ["not_converged"] and s.learned_iterations == 45
    assert s.converged and s.method == "jacobi_pcg"


def test_nonsymmetric_matrix_falls_back_to_ilu_gmres(spk):
    A = spk.convdiff_3d(16, 14, 12, 1.0, 0.4, 0.2, 0.1)
    n = A.nrows
    b = spk.uniform_pm1(3, n)
    bad = spk.CsrMatrix.from_host(n, n, np.arange(n + 1), np.arange(n),
                                  np.full(n, np.nan))
    x = spk.DeviceVector.zeros(n)
    s = spk.learned_solve(b, x, spk.LearnedConfig(tolerance=1e-10), A,
                          spk.spai(bad))
    assert s.method == "ilu_gmres"
    assert s.reasons == ["bad_values", "bad_diagonal"]
    assert s.bad_values == n and s.bad_diagonal == n
    assert s.converged and spk.true_relres(A, b, x) < 1e-10


def test_no_preconditioner_and_zero_rhs(spk):
    A, b = _aniso(spk, side=32)
    x = spk.DeviceVector.zeros(A.nrows)
    s = spk.learned_solve(b, x, spk.LearnedConfig(), A, None)
    assert s.reasons == ["no_preconditioner"] and s.method == "jacobi_pcg"
    assert s.converged
    z = spk.DeviceVector.zeros(A.nrows)
    x = spk.DeviceVector.from_numpy(np.ones(A.nrows))
    s = spk.learned_solve(z, x, spk.LearnedConfig(), A, None)
    assert s.converged and s.relres == 0.0 and not np.any(x.numpy())


def test_rejects_bad_configuration(spk):
    A, b = _aniso(spk, side=16)
    x = spk.DeviceVector.zeros(A.nrows)
    for kw in ({"stall_factor": 0.0}, {"stall_factor": 1.5},
               {"stall_cycles": 0}, {"restart": 0}, {"tolerance": 0.0},
               {"fallback": 7}, {"max_iterations": -1}):
        with pytest.raises(spk.InvalidArgument):
            spk.learned_solve(b, x, spk.LearnedConfig(**kw), A, None)
    with pytest.raises(spk.InvalidArgument):  # M of another size
        spk.learned_solve(b, x, spk.LearnedConfig(), A,
                          spk.SpaiModel(1).preconditioner(spk.aniso_q1_2d(8, 1)))
    with pytest.raises(spk.InvalidArgument):  # not an explicit inverse
        spk.learned_solve(b, x, spk.LearnedConfig(), A, spk.jacobi(A))
