"""Learned sparse approximate inverse on the GPU: the device generator and
features are bit-identical to the oracle; the bf16 MFMA forward pass matches
the oracle's bf16 forward; the device loss and gradient match the oracle's
(which is itself pinned by finite differences in test_oracle_spai.py);
training lowers the loss and the trained M takes fewer FGMRES iterations
than the untrained (Jacobi) one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_params(seed, scale3=0.1):
    rng = np.random.default_rng(seed)
    p = []
    for name, shp in [("W1", (24, 64)), ("b1", (64,)), ("W2", (64, 64)),
                      ("b2", (64,)), ("W3", (64, 9)), ("b3", (9,))]:
        s = np.sqrt(6.0 / shp[0]) if len(shp) == 2 else 0.1
        if name in ("W3", "b3"):
            s = scale3
        p.append(rng.uniform(-s, s, size=shp).ravel())
    return np.concatenate(p).astype(np.float32)


@pytest.mark.parametrize("N,seed,sigma,iso", [(2, 0, 1.5, False),
                                              (7, 3, 1.5, False),
                                              (13, 5, 0.0, True),
                                              (64, 11, 1.5, False)])
def test_generator_bit_exact(O, spk, N, seed, sigma, iso):
    A = spk.aniso_q1_2d(N, seed, sigma, iso)
    R = O.aniso_q1_2d(N, seed, sigma, iso)
    rp, ci, v = A.download()
    assert np.array_equal(rp, R.row_ptr)
    assert np.array_equal(ci, R.col_ind)
    assert np.array_equal(v.view(np.uint64), R.vals.view(np.uint64))


@pytest.mark.parametrize("N", [5, 13, 64])
def test_features_bit_exact(O, spk, N):
    A = spk.aniso_q1_2d(N, 21, 1.5)
    X = spk.spai_features(A)
    Xr = O.spai_features(O.aniso_q1_2d(N, 21, 1.5))
    assert np.array_equal(X.view(np.uint32), Xr.view(np.uint32))


def test_untrained_model_is_jacobi(O, spk):
    A = spk.aniso_q1_2d(20, 4, 1.5)
    R = O.aniso_q1_2d(20, 4, 1.5)
    M = spk.SpaiModel(7).predict(A)
    _, _, mv = M.download()
    rows = np.repeat(np.arange(R.nrows), np.diff(R.row_ptr))
    d = R.vals[R.col_ind == rows]
    want = np.where(R.col_ind == rows, 1.0 / d[rows], 0.0)
    assert np.array_equal(mv, want)


@pytest.mark.parametrize("N", [13, 64])
def test_forward_matches_oracle(O, spk, N):
    p = rand_params(5)
    model = spk.SpaiModel(0)
    model.params = p
    assert np.array_equal(model.params, p)
    A = spk.aniso_q1_2d(N, 8, 1.5)
    R = O.aniso_q1_2d(N, 8, 1.5)
    _, _, mv = model.predict(A).download()
    mr = O.spai_predict(p, R)
    rows = np.repeat(np.arange(R.nrows), np.diff(R.row_ptr))
    d = R.vals[R.col_ind == rows][rows]
    # compare the network outputs y = m * a_ii (minus 1 on the diagonal):
    # only the fp32 accumulation order differs, which may flip the bf16
    # rounding of a hidden unit
    err = np.abs((mv - mr) * d)
    assert err.max() < 2e-2 * max(1.0, np.abs(mr * d).max())
    assert np.median(err) < 1e-4


@pytest.mark.parametrize("N", [13, 64])
def test_loss_and_gradient_match_oracle(O, spk, N):
    p = rand_params(6, scale3=0.2)
    model = spk.SpaiModel(0)
    model.params = p
    A = spk.aniso_q1_2d(N, 9, 1.5)
    R = O.aniso_q1_2d(N, 9, 1.5)
    loss, g = model.loss_grad(A)
    lr, gr = O.spai_loss_grad(p, R)
    assert abs(loss - lr) <= 1e-3 * lr
    for name, sl in [("W1", slice(0, 1536)), ("b1", slice(1536, 1600)),
                     ("W2", slice(1600, 5696)), ("b2", slice(5696, 5760)),
                     ("W3", slice(5760, 6336)), ("b3", slice(6336, 6345))]:
        d = np.linalg.norm(g[sl] - gr[sl])
        assert d <= 2e-2 * np.linalg.norm(gr[sl]) + 1e-7, name
    # deterministic: a second evaluation is bit-identical
    loss2, g2 = model.loss_grad(A)
    assert loss2 == loss and np.array_equal(g2, g)


def test_rejects_rows_beyond_slots(spk):
    A = spk.powerlaw_spd(2000, 3)
    model = spk.SpaiModel(0)
    with pytest.raises(spk.InvalidArgument):
        model.predict(A)
    with pytest.raises(spk.InvalidArgument):
        model.loss_grad(A)


def test_training_helps_fgmres(O, spk):
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(24)]
    model = spk.SpaiModel(1)
    hist = model.train(mats, epochs=12, learning_rate=2e-3, shuffle_seed=5)
    assert len(hist) == 12 * 24
    assert np.all(np.isfinite(hist))
    assert hist[-24:].mean() < 0.8 * hist[:24].mean()
    jac = spk.SpaiModel(1)  # untrained: the Jacobi inverse
    gains = []
    for seed in (1000, 1001):
        A = spk.aniso_q1_2d(64, seed, 1.5)
        b = spk.uniform_pm1(7, A.nrows)
        its = []
        for m in (jac, model):
            x = spk.DeviceVector.zeros(A.nrows)
            s = spk.fgmres(b, x, spk.FgmresConfig(1e-8, 30, 20000), A,
                           m.preconditioner(A))
            assert s.converged
            rel = spk.true_relres(A, b, x)
            assert rel < 1e-7
            its.append(s.computed_iterations)
        gains.append(its[0] / its[1])
        # the trained network's loss on this unseen matrix is lower too
        assert model.loss_grad(A)[0] < jac.loss_grad(A)[0]
    assert min(gains) > 1.1, gains


# ---- the symmetric objective (1/n) ||I - S A||_F^2, S = (M + M^T) / 2
# (spk_spai_model_set_objective; no reference counterpart)


@pytest.mark.parametrize("N", [13, 64])
def test_symmetric_loss_and_gradient_match_oracle(O, spk, N):
    p = rand_params(6, scale3=0.2)
    model = spk.SpaiModel(0)
    model.params = p
    model.objective = spk.SPAI_LOSS_SYMMETRIC
    assert model.objective == spk.SPAI_LOSS_SYMMETRIC
    A = spk.aniso_q1_2d(N, 9, 1.5)
    R = O.aniso_q1_2d(N, 9, 1.5)
    loss, g = model.loss_grad(A)
    lr, gr = O.spai_loss_grad(p, R, symmetric=True)
    assert abs(loss - lr) <= 1e-3 * lr
    for name, sl in [("W1", slice(0, 1536)), ("b1", slice(1536, 1600)),
                     ("W2", slice(1600, 5696)), ("b2", slice(5696, 5760)),
                     ("W3", slice(5760, 6336)), ("b3", slice(6336, 6345))]:
        d = np.linalg.norm(g[sl] - gr[sl])
        assert d <= 2e-2 * np.linalg.norm(gr[sl]) + 1e-7, name
    loss2, g2 = model.loss_grad(A)
    assert loss2 == loss and np.array_equal(g2, g)
    # the other objective on the same model gives the plain loss again
    model.objective = spk.SPAI_LOSS_INVERSE
    assert abs(model.loss_grad(A)[0] - O.spai_loss_grad(p, R)[0]) <= 1e-3 * lr


def test_symmetric_objective_rejects_nonsymmetric_pattern(O, spk):
    A = spk.CsrMatrix.from_oracle(
        O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 1, 2], [2.0, -1.0, 2.0, 2.0]))
    model = spk.SpaiModel(0)
    model.objective = spk.SPAI_LOSS_SYMMETRIC
    with pytest.raises(spk.InvalidArgument, match="structurally symmetric"):
        model.loss_grad(A)
    with pytest.raises(spk.InvalidArgument, match="structurally symmetric"):
        model.train([A], epochs=1)
    with pytest.raises(spk.InvalidArgument):
        model.objective = 7


def test_symmetric_objective_trains_the_symmetric_part(spk):
    """Trained on S's own loss, the network reaches a lower S-loss than the
    same training on the plain loss, and its S preconditions CG."""
    mats = [spk.aniso_q1_2d(48, s, 1.5) for s in range(16)]
    plain, sym = spk.SpaiModel(1), spk.SpaiModel(1)
    sym.objective = spk.SPAI_LOSS_SYMMETRIC
    for m in (plain, sym):
        h = m.train(mats, epochs=6, learning_rate=2e-3, shuffle_seed=5)
        assert np.all(np.isfinite(h))
    A = spk.aniso_q1_2d(64, 1000, 1.5)
    plain.objective = spk.SPAI_LOSS_SYMMETRIC
    assert sym.loss_grad(A)[0] < plain.loss_grad(A)[0]
    b = spk.uniform_pm1(7, A.nrows)
    its = []
    for M in (spk.jacobi(A), spk.spai(sym.predict(A).symmetric_part())):
        x = spk.DeviceVector.zeros(A.nrows)
        s = spk.conjugate_gradient(b, x, spk.CGConfig(1e-8, 20000, 20000), A, M)
        assert s.converged and spk.true_relres(A, b, x) < 1e-7
        its.append(s.computed_iterations)
    assert its[1] < 0.8 * its[0], its
