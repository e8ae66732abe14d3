"""Device permutations against the oracle's restatement of
permutation.hpp:41-148 (itself pinned by dense permutation matrices and the
reference's known answers): identical integer arrays and values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["rows", "cols", "both"])
@pytest.mark.parametrize("which", ["kat", "convdiff", "powerlaw"])
def test_bit_exact(O, spk, mode, which):
    if which == "kat":
        A = O.make_matrix(3, 3, [0, 0, 1, 1, 2, 2], [0, 1, 1, 2, 0, 2],
                          [1., 2., 3., 4., 5., 6.])
        perm = [2, 0, 1]
    elif which == "convdiff":
        A = O.convdiff_3d(9, 8, 7, 1.0, 0.3, 0.2, 0.1)
        perm = np.random.default_rng(1).permutation(A.nrows)
    else:
        A = O.powerlaw_spd(5000, 2)
        perm = np.random.default_rng(2).permutation(A.nrows)
    B = spk.CsrMatrix.from_oracle(A).permute(perm, mode)
    rp, ci, v = B.download()
    R = O.permute(A, perm, mode)
    assert np.array_equal(rp, R.row_ptr)
    assert np.array_equal(ci, R.col_ind)
    assert np.array_equal(v.view(np.uint64), R.vals.view(np.uint64))


def test_rectangular_and_errors(O, spk):
    A = O.make_matrix(3, 5, [0, 0, 1, 2, 2], [1, 4, 2, 0, 3], [1., 2., 3., 4., 5.])
    Ad = spk.CsrMatrix.from_oracle(A)
    for mode, perm in (("rows", [2, 0, 1]), ("cols", [4, 3, 2, 1, 0])):
        rp, ci, v = Ad.permute(perm, mode).download()
        R = O.permute(A, perm, mode)
        assert np.array_equal(ci, R.col_ind) and np.array_equal(v, R.vals)
    with pytest.raises(spk.InvalidArgument):
        Ad.permute([0, 0, 1], "rows")
    with pytest.raises(spk.InvalidArgument):
        Ad.permute([0, 1, 5], "rows")
    with pytest.raises(spk.InvalidArgument):
        Ad.permute([0, 1, 2], "both")
