"""The oracle's symmetric part (M + M^T)/2 on a structurally symmetric
pattern (no reference counterpart), pinned against the dense formula, exact
symmetry, idempotence and the pattern check."""
import numpy as np
import pytest


def test_matches_dense_and_is_exactly_symmetric(O):
    A = O.convdiff_centered_2d(7, 6, 1.0, 0.3, 0.2, 1.0)  # nonsymmetric values
    S = O.symmetric_part(A)
    D = A.dense()
    assert np.allclose(S.dense(), 0.5 * (D + D.T), rtol=0, atol=1e-15)
    Sd = S.dense()
    assert np.array_equal(Sd, Sd.T)
    assert np.array_equal(S.row_ptr, A.row_ptr)
    assert np.array_equal(S.col_ind, A.col_ind)
    S2 = O.symmetric_part(S)
    assert np.array_equal(S2.vals, S.vals)
    assert not np.allclose(S.vals, A.vals)


def test_symmetric_input_unchanged(O):
    A = O.poisson_2d(6, 5, 0.1)
    assert np.array_equal(O.symmetric_part(A).vals, A.vals)


def test_pattern_must_be_symmetric(O):
    M = O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 1, 2], [1., 2., 3., 4.])
    with pytest.raises(O.OracleError):
        O.symmetric_part(M)
