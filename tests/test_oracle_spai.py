"""The oracle's learned sparse approximate inverse, pinned independently:
features against a numpy restatement, the loss against a dense
||I - M A||_F^2, the float64 gradient against central finite differences,
the bf16 gradient against the float64 one, zero output weights against the
Jacobi preconditioner, and bf16 rounding against exhaustive properties."""
import numpy as np
import pytest


def rand_params(seed, scale3=0.05):
    rng = np.random.default_rng(seed)
    p = []
    for name, shp in [("W1", (24, 64)), ("b1", (64,)), ("W2", (64, 64)),
                      ("b2", (64,)), ("W3", (64, 9)), ("b3", (9,))]:
        fan = shp[0] if len(shp) == 2 else 64
        s = np.sqrt(6.0 / fan) if len(shp) == 2 else 0.1
        if name in ("W3", "b3"):
            s = scale3
        p.append(rng.uniform(-s, s, size=shp).ravel())
    return np.concatenate(p)


def slots(A, i):
    rp, ci = A.row_ptr, A.col_ind
    cols = list(ci[rp[i]:rp[i + 1]])
    d = cols.index(i)
    pos = [-1] * 9
    pos[4] = rp[i] + d
    for k in range(d):
        pos[3 - k] = rp[i] + d - 1 - k
    for k in range(len(cols) - d - 1):
        pos[5 + k] = rp[i] + d + 1 + k
    return pos


def np_features(A):
    n = A.nrows
    D = np.array([A.at(i, i) for i in range(n)])
    rs = np.array([sum(A.vals[A.row_ptr[i]:A.row_ptr[i + 1]]) for i in range(n)])
    X = np.zeros((n, 24))
    for i in range(n):
        pos = slots(A, i)
        for k, s in enumerate([0, 1, 2, 3, 5, 6, 7, 8]):
            if pos[s] < 0:
                continue
            j = A.col_ind[pos[s]]
            X[i, k] = A.vals[pos[s]] / D[i]
            X[i, 8 + k] = (D[j] - D[i]) / (D[j] + D[i])
            X[i, 16 + k] = rs[j] / D[j]
    return X.astype(np.float32)


@pytest.fixture(scope="module")
def small(O):
    return O.aniso_q1_2d(7, 3, 1.5)


def test_features_match_numpy(O, small):
    X = O.spai_features(small)
    assert np.array_equal(X, np_features(small))


def test_zero_output_layer_is_jacobi(O, small):
    p = rand_params(1)
    u = O.spai_unpack(p)
    u["W3"][:] = 0
    u["b3"][:] = 0
    m = O.spai_predict(p, small, exact=True)
    D = small.dense()
    want = np.where(small.col_ind == np.repeat(np.arange(small.nrows),
                                               np.diff(small.row_ptr)),
                    1.0 / np.repeat(np.diag(D), np.diff(small.row_ptr)), 0.0)
    assert np.array_equal(m, want)
    loss, _ = O.spai_loss_grad(p, small, exact=True)
    R = np.eye(small.nrows) - np.diag(1 / np.diag(D)) @ D
    assert abs(loss - (R ** 2).sum() / small.nrows) < 1e-12 * max(loss, 1)


@pytest.mark.parametrize("exact", [True, False])
def test_loss_matches_dense(O, small, exact):
    p = rand_params(2)
    m = O.spai_predict(p, small, exact=exact)
    M = small.__class__(small.nrows, small.ncols, small.row_ptr,
                        small.col_ind, m).dense()
    R = np.eye(small.nrows) - M @ small.dense()
    loss, _ = O.spai_loss_grad(p, small, exact=exact)
    assert abs(loss - (R ** 2).sum() / small.nrows) <= 1e-11 * loss


def test_exact_gradient_matches_finite_differences(O, small):
    p = rand_params(3, scale3=0.2)
    _, g = O.spai_loss_grad(p, small, exact=True)
    rng = np.random.default_rng(0)
    # a sample from every parameter block, plus the largest gradients
    idx = list(rng.choice(6345, 40, replace=False)) + list(np.argsort(-abs(g))[:10])
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.spai_loss_grad(pp, small, exact=True)[0] -
              O.spai_loss_grad(pm, small, exact=True)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * np.abs(g).max() + 1e-6 * abs(g[k]), k


def test_bf16_gradient_close_to_exact(O, small):
    p = rand_params(4, scale3=0.2)
    lb, gb = O.spai_loss_grad(p.astype(np.float32), small)
    le, ge = O.spai_loss_grad(p, small, exact=True)
    assert abs(lb - le) < 2e-2 * le
    assert np.linalg.norm(gb - ge) < 5e-2 * np.linalg.norm(ge)


def test_bf16_rounding(O):
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 0.0,
                  65504.0, 1e-30, np.inf, -np.inf], np.float32)
    r = O.bf16_round(x)
    # exactly representable values are kept; ties go to even
    assert r[0] == 1.0 and r[1] == 1.0 and r[3] == 1.0078125
    assert r[2] == 1.0078125
    bits = r.view(np.uint32)
    assert np.all((bits & 0xFFFF) == 0)
    # nearest: |r - x| <= half an ulp of bf16
    fin = np.isfinite(x) & (x != 0)
    ulp = 2.0 ** (np.floor(np.log2(np.abs(x[fin]))) - 7)
    assert np.all(np.abs(r[fin].astype(np.float64) - x[fin]) <= ulp / 2)
    assert np.isinf(r[8]) and np.isinf(r[9])


def test_rejects_rows_beyond_slots(O):
    n = 7
    rows = [0] * 7 + list(range(1, n))
    cols = list(range(7)) + list(range(1, n))
    # row 0 holds 6 entries right of its diagonal
    A = O.make_matrix(n, n, rows, cols, [4.0] + [-0.1] * 6 + [1.0] * (n - 1))
    with pytest.raises(O.OracleError):
        O.spai_features(A)


# ---- the symmetric objective (1/n) ||I - S A||_F^2, S = (M + M^T) / 2 (no
# reference counterpart), pinned like the plain one


@pytest.mark.parametrize("exact", [True, False])
def test_symmetric_loss_matches_dense(O, small, exact):
    p = rand_params(5, scale3=0.2)
    m = O.spai_predict(p, small, exact=exact)
    M = small.__class__(small.nrows, small.ncols, small.row_ptr,
                        small.col_ind, m).dense()
    S = 0.5 * (M + M.T)
    R = np.eye(small.nrows) - S @ small.dense()
    loss, _ = O.spai_loss_grad(p, small, exact=exact, symmetric=True)
    assert abs(loss - (R ** 2).sum() / small.nrows) <= 1e-11 * loss
    # a genuinely nonsymmetric M: the two objectives differ
    assert abs(loss - O.spai_loss_grad(p, small, exact=exact)[0]) > 1e-6 * loss


def test_symmetric_gradient_matches_finite_differences(O, small):
    p = rand_params(6, scale3=0.2)
    _, g = O.spai_loss_grad(p, small, exact=True, symmetric=True)
    rng = np.random.default_rng(1)
    idx = list(rng.choice(6345, 40, replace=False)) + list(np.argsort(-abs(g))[:10])
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.spai_loss_grad(pp, small, exact=True, symmetric=True)[0] -
              O.spai_loss_grad(pm, small, exact=True, symmetric=True)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * np.abs(g).max() + 1e-6 * abs(g[k]), k


def test_symmetric_loss_equals_plain_for_symmetric_m(O, small):
    """A zero output layer predicts M = D^-1, already symmetric: S = M bit
    for bit, so the two losses agree exactly (their gradients do not: the
    symmetric one averages each entry's with its mirror's)."""
    p = rand_params(7)
    u = O.spai_unpack(p)
    u["W3"][:] = 0
    u["b3"][:] = 0
    for exact in (True, False):
        a = O.spai_loss_grad(p, small, exact=exact)
        b = O.spai_loss_grad(p, small, exact=exact, symmetric=True)
        assert a[0] == b[0]
        assert not np.array_equal(a[1], b[1])


def test_symmetric_loss_needs_symmetric_pattern(O):
    # 3 x 3, diagonal plus one entry above it: (0, 1) has no mirror
    A = O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 1, 2], [2.0, -1.0, 2.0, 2.0])
    p = rand_params(8)
    assert np.isnan(O.spai_loss_grad(p, A, exact=True, symmetric=True)[0])
    assert np.isfinite(O.spai_loss_grad(p, A, exact=True)[0])
