"""The oracle's reverse Cuthill-McKee (reordering.cpp:19-203), pinned by the
reference's test cases (reordering_test.cpp:42-300: degrees, the path
graph's pseudo-peripheral node, valid permutations, bandwidth never grows,
disconnected and diagonal patterns, values travel with dperm) and by an
independent check of the Cuthill-McKee order itself: every node after the
first of its component is reached from the earliest ordered neighbour, and
consecutive nodes with the same earliest parent appear by nondecreasing
degree, ties by index."""
import numpy as np
import pytest


def _pattern(O, n, pairs):
    r = [i for i, j in pairs]
    c = [j for i, j in pairs]
    return O.make_matrix(n, n, r, c, [1.0] * len(pairs))


def _bandwidth(O, A):
    lo = hi = 0
    for i in range(A.nrows):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            j = int(A.col_ind[k])
            lo, hi = max(lo, i - j), max(hi, j - i)
    return lo, hi


PATH4 = [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2), (2, 1), (2, 2), (2, 3), (3, 2),
         (3, 3)]
ARROW5 = [(0, 0), (0, 1), (0, 2), (0, 3), (0, 4), (1, 0), (1, 1), (2, 0), (2, 2),
          (3, 0), (3, 3), (4, 0), (4, 4)]


def test_reference_cases(O):
    P3 = _pattern(O, 3, [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2), (2, 1), (2, 2)])
    rp, ci = O.symmetrize_pattern(P3)
    assert list(O.adjacency_degree(rp, ci)) == [1, 2, 1]
    assert O.pseudo_peripheral_node(_pattern(O, 4, PATH4)) in (0, 3)
    for n, pairs in ((4, PATH4), (5, ARROW5),
                     (4, [(0, 0), (0, 3), (1, 1), (1, 2), (2, 2), (2, 3), (3, 0),
                          (3, 3)]),
                     (4, [(0, 0), (0, 1), (1, 0), (1, 1), (2, 2), (2, 3), (3, 2),
                          (3, 3)]),
                     (2, [(0, 0), (1, 1)])):
        A = _pattern(O, n, pairs)
        perm = O.reverse_cuthill_mckee(A)
        assert O.is_valid_permutation(perm)
        B = O.permute(A, perm, "both")
        lo0, hi0 = _bandwidth(O, A)
        lo1, hi1 = _bandwidth(O, B)
        assert max(lo1, hi1) <= max(lo0, hi0)
        assert B.nnz == A.nnz


def _check_cm_order(O, A, perm):
    """Independent of the restatement: the reversed permutation is a
    Cuthill-McKee order of the symmetrised graph."""
    rp, ci = O.symmetrize_pattern(A)
    deg = O.adjacency_degree(rp, ci)
    n = A.nrows
    order = np.empty(n, np.int64)
    order[n - 1 - perm] = np.arange(n)
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    parent = np.full(n, -1)
    for k, v in enumerate(order):
        nbrs = [int(ci[t]) for t in range(rp[v], rp[v + 1]) if ci[t] != v]
        earlier = [pos[u] for u in nbrs if pos[u] < k]
        parent[k] = min(earlier) if earlier else -1
    for k in range(1, n):
        if parent[k] == -1:
            continue  # a component's start
        # the earliest parent of consecutive nodes never decreases ...
        prev = k - 1
        if parent[prev] != -1:
            assert parent[prev] <= parent[k]
            if parent[prev] == parent[k]:  # ... and within it, (deg, index)
                a, b = order[prev], order[k]
                assert (deg[a], a) < (deg[b], b)


@pytest.mark.parametrize("make", [
    lambda O: O.poisson_2d(7, 5, 0.1),
    lambda O: O.convdiff_3d(4, 3, 5, 1.0, 0.3, 0.2, 0.1),
    lambda O: O.aniso_q1_2d(6, 2, 1.5),
    lambda O: O.nonsymmetric_sample()])
def test_cuthill_mckee_property(O, make):
    A = make(O)
    perm = O.reverse_cuthill_mckee(A)
    assert O.is_valid_permutation(perm)
    _check_cm_order(O, A, perm)


def test_scrambled_band_recovered(O):
    """A path graph under a random relabelling: RCM restores bandwidth 1."""
    n = 40
    rng = np.random.default_rng(3)
    lab = rng.permutation(n)
    pairs = [(lab[i], lab[i]) for i in range(n)]
    pairs += [(lab[i], lab[i + 1]) for i in range(n - 1)]
    pairs += [(lab[i + 1], lab[i]) for i in range(n - 1)]
    A = _pattern(O, n, pairs)
    B = O.permute(A, O.reverse_cuthill_mckee(A), "both")
    assert _bandwidth(O, B) == (1, 1)
