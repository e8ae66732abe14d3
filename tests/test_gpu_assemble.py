"""Device CSR assembly from (row, col[, value]) lists against the oracle's
restatement of Compressed_row_sparsity (Compressed_row_sparsity_Impl.cpp:
10-40) and make_matrix (matgen.hpp:25-52): identical integer arrays,
last-duplicate-wins values, the reference's error conventions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_reference_kats(spk):
    # compressed_data_test.cpp:142-203
    rp, ci, _ = spk.CsrMatrix.assemble(4, 6, [2, 2, 2], [5, 2, 4]).download()
    assert list(ci) == [2, 4, 5]
    A = spk.CsrMatrix.assemble(4, 5, [2, 2, 3], [3, 3, 4])
    rp, ci, _ = A.download()
    assert A.nnz == 2 and rp[3] == 1 and rp[4] == 2
    rp, ci, _ = spk.CsrMatrix.assemble(3, 3, [], []).download()
    assert list(rp) == [0, 0, 0, 0] and len(ci) == 0
    rp, ci, _ = spk.CsrMatrix.assemble(4, 8, [2, 2, 2, 2], [7, 3, 2, 5]).download()
    assert list(rp) == [0, 0, 0, 4, 4] and list(ci) == [2, 3, 5, 7]


@pytest.mark.parametrize("seed,n,m,k", [(0, 7, 5, 30), (1, 100, 80, 2000),
                                        (2, 3000, 2500, 200000)])
def test_random_lists_bit_exact(O, spk, seed, n, m, k):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, k)
    cols = rng.integers(0, m, k)
    # plenty of repeated positions with different values
    rows[k // 2:] = rows[:k - k // 2]
    cols[k // 2:] = cols[:k - k // 2]
    vals = rng.standard_normal(k)
    A = spk.CsrMatrix.assemble(n, m, rows, cols, vals)
    rp, ci, v = A.download()
    R = O.make_matrix(n, m, rows, cols, vals)
    assert np.array_equal(rp, R.row_ptr)
    assert np.array_equal(ci, R.col_ind)
    assert np.array_equal(v.view(np.uint64), R.vals.view(np.uint64))


def test_empty_rows(spk):
    A = spk.CsrMatrix.assemble(5, 4, [3], [1], [2.5])
    rp, ci, v = A.download()
    assert list(rp) == [0, 0, 0, 0, 1, 1] and list(ci) == [1] and v[0] == 2.5


def test_errors(spk):
    with pytest.raises(spk.LogicError):          # Shape.cpp:8-10
        spk.CsrMatrix.assemble(1, 4, [0], [0])
    with pytest.raises(spk.LogicError):          # Index.cpp:8
        spk.CsrMatrix.assemble(4, 4, [0, -1], [0, 0])
    with pytest.raises(spk.InvalidArgument):     # outside the shape
        spk.CsrMatrix.assemble(4, 4, [0, 4], [0, 0])
