"""The oracle's batched variable-coefficient diffusion family, pinned by the
energy identity x^T A x = sum over faces of t (x_a - x_b)^2 + sum over
boundary faces of t x_a^2 (independent of the assembly order), bitwise
symmetry, the M-matrix sign pattern, closed-form nnz and the laws of the
side lengths and conductivities."""
import numpy as np
import pytest


@pytest.mark.parametrize("k", [0, 3, 17])
def test_energy_identity(O, k):
    seed, smin, smax = 11, 5, 12
    A = O.batched_diffusion(seed, k, smin, smax)
    kap = O.batched_kappa(seed, k, smin, smax)
    s = kap.shape[0]
    rng = np.random.default_rng(k)
    x = rng.standard_normal(s * s)
    X = x.reshape(s, s)
    e = 0.0
    for y in range(s):
        for xx in range(s):
            kc = kap[y, xx]
            for dy, dx in ((0, 1), (1, 0)):
                yy, x2 = y + dy, xx + dx
                if yy < s and x2 < s:
                    t = 2 * kc * kap[yy, x2] / (kc + kap[yy, x2])
                    e += t * (X[y, xx] - X[yy, x2]) ** 2
            nb = (y == 0) + (y == s - 1) + (xx == 0) + (xx == s - 1)
            e += nb * 2 * kc * X[y, xx] ** 2
    got = x @ O.spmv(A, x)
    assert abs(got - e) <= 1e-12 * e
    assert A.nnz == 5 * s * s - 4 * s


def test_symmetric_m_matrix(O):
    A = O.batched_diffusion(3, 1, 6, 9)
    D = A.dense()
    assert np.array_equal(D, D.T)
    off = D - np.diag(np.diag(D))
    assert np.all(off <= 0)
    d, rs = np.diag(D), -off.sum(axis=1)
    assert np.all(d >= rs * (1 - 1e-14))        # weakly dominant
    s = int(round(np.sqrt(D.shape[0])))
    edge = np.zeros((s, s), bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    assert np.all(d[edge.ravel()] > rs[edge.ravel()])  # Dirichlet rows strictly
    assert np.linalg.eigvalsh(D).min() > 0


def test_laws(O):
    sides = np.array([O.batched_side(42, k) for k in range(4096)])
    assert sides.min() == 32 and sides.max() == 256
    assert abs(sides.mean() - 144) < 3
    kap = O.batched_kappa(42, 7)
    lk = np.log(kap)
    assert abs(lk.mean()) < 0.05 and abs(lk.std() - 1.0) < 0.05
