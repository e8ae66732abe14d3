"""The oracle's permutations (permutation.cpp:15-35, permutation.hpp:41-148),
pinned by dense permutation matrices (rperm = P A, cperm = A P^T,
dperm = P A P^T with P[perm[i], i] = 1), the reference's own known answers
(permutation_test.cpp) and validity rules."""
import numpy as np
import pytest


def pmat(perm):
    n = len(perm)
    P = np.zeros((n, n))
    P[np.asarray(perm), np.arange(n)] = 1.0
    return P


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_against_dense(O, seed):
    rng = np.random.default_rng(seed)
    A = O.convdiff_centered_2d(5, 4, 1.0, 0.3, 0.2, 3.0)
    n = A.nrows
    perm = rng.permutation(n)
    D, P = A.dense(), pmat(perm)
    for mode, want in [("rows", P @ D), ("cols", D @ P.T),
                       ("both", P @ D @ P.T)]:
        B = O.permute(A, perm, mode)
        assert np.array_equal(B.dense(), want), mode
        for r in range(n):  # sorted columns in every row
            assert np.all(np.diff(B.col_ind[B.row_ptr[r]:B.row_ptr[r + 1]]) > 0)
        assert B.nnz == A.nnz


def test_identity_and_inverse(O):
    A = O.poisson_2d(4, 3, 0.0)
    n = A.nrows
    B = O.permute(A, np.arange(n), "both")
    assert np.array_equal(B.row_ptr, A.row_ptr)
    assert np.array_equal(B.col_ind, A.col_ind)
    assert np.array_equal(B.vals, A.vals)
    perm = np.random.default_rng(3).permutation(n)
    inv = np.argsort(perm)
    C = O.permute(O.permute(A, perm, "both"), inv, "both")
    assert np.array_equal(C.col_ind, A.col_ind) and np.array_equal(C.vals, A.vals)


def test_rectangular_and_validity(O):
    A = O.make_matrix(3, 5, [0, 0, 1, 2, 2], [1, 4, 2, 0, 3],
                      [1., 2., 3., 4., 5.])
    B = O.permute(A, [2, 0, 1], "rows")
    assert np.array_equal(B.dense(), pmat([2, 0, 1]) @ A.dense())
    C = O.permute(A, [4, 3, 2, 1, 0], "cols")
    assert np.array_equal(C.dense(), A.dense() @ pmat([4, 3, 2, 1, 0]).T)
    assert O.is_valid_permutation([2, 0, 1])
    assert not O.is_valid_permutation([0, 0, 1])
    assert not O.is_valid_permutation([0, 3, 1])
    assert not O.is_valid_permutation([0, -1, 1])
    with pytest.raises(O.OracleError):
        O.permute(A, [0, 0, 1], "rows")
    with pytest.raises(O.OracleError):
        O.permute(A, [0, 1, 2], "both")   # not square


def kat_matrix(O):
    # permutation_test.cpp:119-156: A = [[1,2,0],[0,3,4],[5,0,6]]
    return O.make_matrix(3, 3, [0, 0, 1, 1, 2, 2], [0, 1, 1, 2, 0, 2],
                         [1., 2., 3., 4., 5., 6.])


def test_reference_kats(O):
    A = kat_matrix(O)
    B = O.permute(A, [2, 0, 1], "rows").dense()     # permutation_test.cpp:119-154
    assert np.array_equal(B, [[0, 3, 4], [5, 0, 6], [1, 2, 0]])
    C = O.permute(A, [2, 0, 1], "cols").dense()     # permutation_test.cpp:211-249
    assert np.array_equal(C, [[2, 0, 1], [3, 4, 0], [0, 6, 5]])
    # dperm equals rperm of cperm (permutation_test.cpp:308)
    D = O.permute(A, [2, 0, 1], "both")
    E = O.permute(O.permute(A, [2, 0, 1], "cols"), [2, 0, 1], "rows")
    assert np.array_equal(D.col_ind, E.col_ind) and np.array_equal(D.vals, E.vals)
