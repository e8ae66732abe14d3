"""The anisotropic Q1 diffusion generator of the oracle, pinned by the
textbook Q1 Laplacian stencil, an element stiffness computed independently
by Gauss quadrature, a brute-force global assembly, bitwise symmetry,
positive definiteness and the coefficient law."""
import math

import numpy as np
import pytest


def tensor_from_seed(O, seed):
    m = (1 << 64) - 1
    u = [(O.splitmix64_at(seed ^ 0x414E49, k) >> 11) / 2.0 ** 53
         for k in range(2)]
    eps = math.exp(math.log(1000.0) * u[0])
    th = math.pi * u[1]
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    return R @ np.diag([1.0, eps]) @ R.T


def gauss_element(K):
    """Element stiffness of the bilinear square element by 2x2 Gauss
    quadrature (exact for these degree-2 integrands); corners q = dx + 2 dy."""
    g = [0.5 - 0.5 / math.sqrt(3.0), 0.5 + 0.5 / math.sqrt(3.0)]
    corners = [(0, 0), (1, 0), (0, 1), (1, 1)]
    E = np.zeros((4, 4))
    for x in g:
        for y in g:
            grads = []
            for (dx, dy) in corners:
                fx = (x if dx else 1 - x)
                fy = (y if dy else 1 - y)
                gx = (1 if dx else -1) * fy
                gy = fx * (1 if dy else -1)
                grads.append(np.array([gx, gy]))
            for a in range(4):
                for b in range(4):
                    E[a, b] += 0.25 * grads[a] @ K @ grads[b]
    return E


def test_isotropic_is_textbook_q1_laplacian(O):
    N = 6
    A = O.aniso_q1_2d(N, 0, sigma=0.0, isotropic=True)
    D = A.dense()
    p = 2 * N + 2   # interior node (3, 3)
    nb = [p + oy * N + ox for oy in (-1, 0, 1) for ox in (-1, 0, 1)]
    for q in nb:
        want = 8.0 / 3.0 if q == p else -1.0 / 3.0
        assert abs(D[p, q] - want) < 1e-15
    assert A.nnz == (3 * N - 2) ** 2


@pytest.mark.parametrize("seed", [0, 1, 17, 511])
def test_element_table_matches_quadrature(O, seed):
    _, E = O.aniso_q1_params(4, seed, 1.5)
    K = tensor_from_seed(O, seed)
    Eg = gauss_element(K)
    assert np.allclose(E, Eg, rtol=0, atol=1e-12 * np.abs(Eg).max())
    assert np.all(E == E.T)
    assert np.allclose(E.sum(axis=1), 0.0, atol=1e-12 * np.abs(E).max())


@pytest.mark.parametrize("N,seed", [(2, 3), (5, 4), (9, 100)])
def test_matches_bruteforce_assembly(O, N, seed):
    kappa, E = O.aniso_q1_params(N, seed, 1.5)
    nc = N + 1
    G = np.zeros((N * N, N * N))
    for cy in range(nc):
        for cx in range(nc):
            nodes = []
            for q in range(4):
                ix, iy = cx + (q & 1), cy + (q >> 1)
                if 1 <= ix <= N and 1 <= iy <= N:
                    nodes.append((q, (iy - 1) * N + ix - 1))
            for qa, pa in nodes:
                for qb, pb in nodes:
                    G[pa, pb] += kappa[cy * nc + cx] * E[qa, qb]
    A = O.aniso_q1_2d(N, seed, 1.5)
    D = A.dense()
    assert np.allclose(D, G, rtol=1e-14, atol=1e-14 * np.abs(G).max())
    # pattern: exactly the 9-point neighbourhood, columns sorted
    for p in range(N * N):
        cols = A.col_ind[A.row_ptr[p]:A.row_ptr[p + 1]]
        ix, iy = p % N, p // N
        want = sorted((iy + oy) * N + ix + ox for oy in (-1, 0, 1)
                      for ox in (-1, 0, 1)
                      if 0 <= ix + ox < N and 0 <= iy + oy < N)
        assert list(cols) == want


@pytest.mark.parametrize("seed", [0, 5])
def test_symmetric_positive_definite(O, seed):
    A = O.aniso_q1_2d(12, seed, 1.5)
    D = A.dense()
    assert np.array_equal(D, D.T)          # bitwise
    assert np.linalg.eigvalsh(D).min() > 0


def test_coefficient_law(O):
    kappa, _ = O.aniso_q1_params(255, 9, 1.5)
    lk = np.log(kappa)
    assert abs(lk.mean()) < 0.02
    assert abs(lk.std() - 1.5) < 0.02
    eps = [np.linalg.eigvalsh(tensor_from_seed(O, s)) for s in range(200)]
    ratio = np.array([e[1] / e[0] for e in eps])
    assert ratio.min() >= 1.0 - 1e-12 and ratio.max() <= 1000.0 * (1 + 1e-12)
    assert 0.35 < np.mean(np.log10(ratio) < 1.5) < 0.65
