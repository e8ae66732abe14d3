"""The oracle's Matrix Market reader/writer (matrix_market.hpp,
matrix_market.cpp), pinned by the reference's own test strings
(matrix_market_test.cpp:33-245) and by scipy.io's independent Matrix Market
reader on generated files."""
import io

import numpy as np
import pytest

KAT_GENERAL = ("%%MatrixMarket matrix coordinate real general\n3 4 3\n"
               "1 2 4.5\n2 3 7.0\n3 1 2.5\n")
KAT_COMMENTS = ("%%MatrixMarket matrix coordinate real general\n"
                "% This is a comment\n% Another comment\n3 3 2\n"
                "1 1 5.0\n3 3 9.0\n")
KAT_SYMMETRIC = ("%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n"
                 "1 1 1.0\n2 1 2.0\n3 1 3.0\n3 3 4.0\n")
KAT_INTEGER = ("%%MatrixMarket matrix coordinate integer general\n3 3 2\n"
               "1 2 7\n3 1 3\n")
KAT_PATTERN = ("%%MatrixMarket matrix coordinate pattern general\n3 3 3\n"
               "1 1\n2 2\n3 3\n")


def test_reference_reading_cases(O):
    A = O.read_matrix_market_text(KAT_GENERAL)
    assert (A.nrows, A.ncols, A.nnz) == (3, 4, 3)
    D = A.dense()
    assert D[0, 1] == 4.5 and D[1, 2] == 7.0 and D[2, 0] == 2.5
    A = O.read_matrix_market_text(KAT_COMMENTS)
    assert A.nnz == 2 and A.dense()[0, 0] == 5.0 and A.dense()[2, 2] == 9.0
    A = O.read_matrix_market_text(KAT_SYMMETRIC)
    assert A.nnz == 6
    D = A.dense()
    assert D[1, 0] == D[0, 1] == 2.0 and D[2, 0] == D[0, 2] == 3.0
    A = O.read_matrix_market_text(KAT_INTEGER)
    assert A.nnz == 2 and A.dense()[0, 1] == 7.0
    A = O.read_matrix_market_text(KAT_PATTERN)
    assert A.nnz == 3 and np.array_equal(A.dense(), np.eye(3))
    A = O.read_matrix_market_text(
        "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 42.0\n")
    assert A.dense()[0, 0] == 42.0 and A.dense()[1, 1] == 0.0


def test_reference_writing_cases(O):
    A = O.make_matrix(3, 4, [0, 1, 2], [1, 2, 0], [4.5, 7.0, 2.5])
    assert O.write_matrix_market_text(A) == (
        "%%MatrixMarket matrix coordinate real general\n3 4 3\n"
        "1 2 4.5\n2 3 7\n3 1 2.5\n")
    E = O.make_matrix(3, 3, [], [], [])
    assert O.write_matrix_market_text(E) == (
        "%%MatrixMarket matrix coordinate real general\n3 3 0\n")
    R = O.make_matrix(4, 5, [0, 0, 1, 2, 3], [0, 3, 1, 4, 2],
                      [1., 2., 3., 4., 5.])
    back = O.read_matrix_market_text(O.write_matrix_market_text(R))
    assert np.array_equal(back.dense(), R.dense())


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n3 3\n",
    "%%MatrixMarket matrix coordinate complex general\n3 3 1\n1 1 1.0 2.0\n",
    "%%NotMatrixMarket matrix coordinate real general\n3 3 0\n",
    "%%MatrixMarket matrix coordinate real unknown\n3 3 0\n", ""])
def test_unsupported_inputs(O, text):
    with pytest.raises(O.OracleError):
        O.read_matrix_market_text(text)


def test_last_entry_wins_and_scipy_agreement(O):
    A = O.read_matrix_market_text(
        "%%MatrixMarket matrix coordinate real general\n3 3 3\n"
        "1 1 1.0\n2 2 2.0\n1 1 9.0\n")
    assert A.nnz == 2 and A.dense()[0, 0] == 9.0
    scipy_io = pytest.importorskip("scipy.io")
    for M in (O.convdiff_centered_2d(7, 6, 1.0, 0.3, 0.2, 1.0),
              O.poisson_2d(5, 4, 0.25), O.aniso_q1_2d(8, 3, 1.5)):
        text = O.write_matrix_market_text(M)
        mine = O.read_matrix_market_text(text)
        ref = scipy_io.mmread(io.StringIO(text)).toarray()
        assert np.array_equal(mine.dense(), ref)
        # %g keeps 6 significant digits
        assert np.allclose(mine.dense(), M.dense(), rtol=1e-5, atol=0)
        assert np.array_equal(mine.row_ptr, M.row_ptr)
        assert np.array_equal(mine.col_ind, M.col_ind)
    S = O.poisson_2d(4, 4, 1.0)
    lower = ["%%MatrixMarket matrix coordinate real symmetric",
             f"{S.nrows} {S.ncols} "
             f"{sum(1 for i in range(S.nrows) for k in range(S.row_ptr[i], S.row_ptr[i + 1]) if S.col_ind[k] <= i)}"]
    for i in range(S.nrows):
        for k in range(S.row_ptr[i], S.row_ptr[i + 1]):
            if S.col_ind[k] <= i:
                lower.append(f"{i + 1} {S.col_ind[k] + 1} {float(S.vals[k])!r}")
    text = "\n".join(lower) + "\n"
    mine = O.read_matrix_market_text(text)
    assert np.array_equal(mine.dense(), S.dense())
    assert np.array_equal(mine.dense(), scipy_io.mmread(io.StringIO(text)).toarray())
