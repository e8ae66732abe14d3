"""Device reverse Cuthill-McKee (reordering.cpp:19-203) against the oracle:
the same permutation exactly (level-synchronous order = the sequential
breadth-first order), the same pseudo-peripheral node, and dperm with the
device permutation narrowing the band."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pattern(O, n, pairs):
    return O.make_matrix(n, n, [i for i, j in pairs], [j for i, j in pairs],
                         [1.0] * len(pairs))


def _cases(O):
    rng = np.random.default_rng(11)
    # two components plus isolated nodes in the middle and at the end
    pairs = [(i, i) for i in range(30)]
    pairs += [(i, i + 1) for i in range(0, 9)] + [(i + 1, i) for i in range(0, 9)]
    pairs += [(i, j) for i in range(15, 25) for j in range(15, 25)
              if abs(i - j) in (1, 3)]
    lab = rng.permutation(50)
    scr = [(lab[i], lab[j]) for i in range(50) for j in range(50)
           if i == j or abs(i - j) == 1 or (i % 7 == 0 and abs(i - j) == 5)]
    return {
        "path4": _pattern(O, 4, [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2), (2, 1),
                                 (2, 2), (2, 3), (3, 2), (3, 3)]),
        "arrow5": _pattern(O, 5, [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (1, 0),
                                  (1, 1), (2, 0), (2, 2), (3, 0), (3, 3), (4, 0),
                                  (4, 4)]),
        "unsym4": _pattern(O, 4, [(0, 0), (0, 3), (1, 1), (1, 2), (2, 2), (2, 3),
                                  (3, 0), (3, 3)]),
        "diag2": _pattern(O, 2, [(0, 0), (1, 1)]),
        "components": _pattern(O, 30, pairs),
        "scrambled": _pattern(O, 50, scr),
        "poisson2d": O.poisson_2d(13, 9, 0.1),
        "convdiff3d": O.convdiff_3d(9, 8, 7, 1.0, 0.3, 0.2, 0.1),
        "aniso_q1": O.aniso_q1_2d(14, 4, 1.5),
        "powerlaw": O.powerlaw_spd(400, 3),
    }


@pytest.mark.parametrize("name", ["path4", "arrow5", "unsym4", "diag2",
                                  "components", "scrambled", "poisson2d",
                                  "convdiff3d", "aniso_q1", "powerlaw"])
def test_rcm_matches_oracle(O, spk, name):
    A = _cases(O)[name]
    d = spk.CsrMatrix.from_oracle(A)
    assert d.pseudo_peripheral_node() == O.pseudo_peripheral_node(A)
    perm = d.rcm()
    ref = O.reverse_cuthill_mckee(A)
    assert np.array_equal(perm, ref)
    B = d.permute(perm, "both")
    rp, ci, v = B.download()
    R = O.permute(A, ref, "both")
    assert np.array_equal(rp, R.row_ptr) and np.array_equal(ci, R.col_ind)
    assert np.array_equal(v, R.vals)


def test_rcm_larger_grid(O, spk):
    """A 3D grid scrambled by a random relabelling: the device order equals
    the oracle's and recovers a narrow band."""
    A0 = O.poisson_3d(12, 11, 10)
    lab = np.random.default_rng(5).permutation(A0.nrows)
    A = O.permute(A0, lab, "both")
    d = spk.CsrMatrix.from_oracle(A)
    perm = d.rcm()
    assert np.array_equal(perm, O.reverse_cuthill_mckee(A))
    rp, ci, _ = d.permute(perm, "both").download()
    band = max(abs(int(ci[k]) - i) for i in range(A.nrows)
               for k in range(rp[i], rp[i + 1]))
    assert band <= 12 * 11 * 2


def test_rcm_rejects_rectangular(O, spk):
    R = O.make_matrix(3, 4, [0, 1, 2], [0, 1, 2], [1., 2., 3.])
    with pytest.raises(spk.InvalidArgument):
        spk.CsrMatrix.from_oracle(R).rcm()
