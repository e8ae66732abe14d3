"""The oracle's power-law SPD generator, pinned by an independent pure-Python
construction from its definition (set union of the draws and their
transposes, the symmetric part, the dominance rule of matgen.hpp:657-658),
the Zipf(1.5) law of the row lengths, bitwise symmetry and definiteness."""
import numpy as np
import pytest

M64 = (1 << 64) - 1


def mix(seed, k):
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def u01(z):
    return (z >> 11) * (1.0 / 9007199254740992.0)


def build(n, seed):
    ks = np.arange(3, 4097, dtype=np.float64)
    cdf = np.cumsum(ks ** -1.5)       # sequential sums, as the definition
    cdf = cdf / cdf[-1]
    drawn = {}
    for i in range(n):
        u = u01(mix(seed ^ 0x4C454E, i))
        L = int(np.searchsorted(cdf, u, side="left")) + 3
        for t in range(L - 1):
            c = mix(seed ^ 0x434F4C, i * 4096 + t) % (n - 1)
            c += c >= i
            drawn.setdefault((i, c), True)
    pat = set(drawn) | {(j, i) for (i, j) in drawn} | {(i, i) for i in range(n)}
    D = np.zeros((n, n))
    for (i, j) in pat:
        if i == j:
            continue
        a = -u01(mix(seed ^ 0x56414C, i * n + j)) if (i, j) in drawn else 0.0
        b = -u01(mix(seed ^ 0x56414C, j * n + i)) if (j, i) in drawn else 0.0
        D[i, j] = 0.5 * (a + b)
    for i in range(n):
        s = 0.0
        for j in range(n):          # ascending j
            if j != i and (i, j) in pat:
                s += abs(D[i, j])
        D[i, i] = s + 1.0
    return pat, D


@pytest.mark.parametrize("n,seed", [(40, 1), (300, 7)])
def test_matches_independent_construction(O, n, seed):
    A = O.powerlaw_spd(n, seed)
    pat, D = build(n, seed)
    got = {(i, int(j)) for i in range(n)
           for j in A.col_ind[A.row_ptr[i]:A.row_ptr[i + 1]]}
    assert got == pat
    for i in range(n):
        cols = A.col_ind[A.row_ptr[i]:A.row_ptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
        assert np.array_equal(A.vals[A.row_ptr[i]:A.row_ptr[i + 1]], D[i, cols])


def test_symmetric_definite(O):
    A = O.powerlaw_spd(400, 3)
    D = A.dense()
    assert np.array_equal(D, D.T)
    off = np.abs(D - np.diag(np.diag(D))).sum(axis=1)
    assert np.all(np.diag(D) > off)
    assert np.linalg.eigvalsh(D).min() > 0


def test_row_length_law(O):
    L = O.powerlaw_lengths(200000, 42)
    assert L.min() >= 3 and L.max() <= 4096
    # Zipf(1.5) on [3, 4096]: mean 101.7, P(L >= 1024) = 2.5 %
    ks = np.arange(3, 4097, dtype=float)
    p = ks ** -1.5 / (ks ** -1.5).sum()
    assert abs(L.mean() - (ks * p).sum()) < 0.03 * (ks * p).sum()
    assert abs(np.mean(L >= 1024) - p[ks >= 1024].sum()) < 0.002
