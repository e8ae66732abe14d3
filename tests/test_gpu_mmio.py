"""Matrix Market text in and out of device stores (matrix_market.hpp):
the store read on the device equals the oracle's to_compressed_row of
read_matrix_market bit for bit, the written text equals the oracle's
(and the reference's test strings), files round-trip, and the reference's
error cases raise the runtime error."""
import os

import numpy as np
import pytest

from test_oracle_mmio import (KAT_COMMENTS, KAT_GENERAL, KAT_INTEGER,
                              KAT_PATTERN, KAT_SYMMETRIC)

pytestmark = pytest.mark.gpu


def _same(spk_store, R):
    rp, ci, v = spk_store.download()
    assert (spk_store.nrows, spk_store.ncols) == (R.nrows, R.ncols)
    assert np.array_equal(rp, R.row_ptr) and np.array_equal(ci, R.col_ind)
    assert np.array_equal(v, R.vals)


@pytest.mark.parametrize("text", [KAT_GENERAL, KAT_COMMENTS, KAT_SYMMETRIC,
                                  KAT_INTEGER, KAT_PATTERN,
                                  "%%MatrixMarket matrix coordinate real general\n"
                                  "3 3 3\n1 1 1.0\n2 2 2.0\n1 1 9.0\n",
                                  "%%matrixmarket MATRIX Coordinate Real General\n"
                                  "2 2 2\n  2   1   -3.25e-2  \n1 2 +7\n"])
def test_read_matches_oracle(O, spk, text):
    _same(spk.CsrMatrix.from_matrix_market_text(text),
          O.read_matrix_market_text(text))


def test_write_matches_oracle_and_round_trips(O, spk, tmp_path):
    A = O.make_matrix(3, 4, [0, 1, 2], [1, 2, 0], [4.5, 7.0, 2.5])
    assert spk.CsrMatrix.from_oracle(A).to_matrix_market_text() == (
        "%%MatrixMarket matrix coordinate real general\n3 4 3\n"
        "1 2 4.5\n2 3 7\n3 1 2.5\n")
    for M in (O.convdiff_centered_2d(9, 7, 1.0, 0.3, 0.2, 1.0),
              O.aniso_q1_2d(12, 4, 1.5), O.powerlaw_spd(300, 5)):
        d = spk.CsrMatrix.from_oracle(M)
        t = d.to_matrix_market_text()
        assert t == O.write_matrix_market_text(M)
        p = os.fspath(tmp_path / "m.mtx")
        d.write_matrix_market(p)
        with open(p) as f:
            assert f.read() == t
        back = spk.CsrMatrix.from_matrix_market(p)
        _same(back, O.read_matrix_market_text(t))
    # a device-generated store, written and read back: same pattern
    G = spk.poisson_3d(10, 9, 8)
    back = spk.CsrMatrix.from_matrix_market_text(G.to_matrix_market_text())
    rp0, ci0, v0 = G.download()
    rp1, ci1, v1 = back.download()
    assert np.array_equal(rp0, rp1) and np.array_equal(ci0, ci1)
    assert np.array_equal(v0, v1)  # integers print exactly with %g


@pytest.mark.parametrize("text,err", [
    ("%%MatrixMarket matrix array real general\n3 3\n", "ReadError"),
    ("%%MatrixMarket matrix coordinate complex general\n3 3 1\n1 1 1.0 2.0\n",
     "ReadError"),
    ("%%NotMatrixMarket matrix coordinate real general\n3 3 0\n", "ReadError"),
    ("", "ReadError"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n",
     "LogicError"),  # a missing entry line reads as index 0 - 1 (Index.cpp:8)
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n",
     "InvalidArgument")])
def test_errors(spk, text, err):
    with pytest.raises(getattr(spk, err)):
        spk.CsrMatrix.from_matrix_market_text(text)


def test_missing_file(spk, tmp_path):
    with pytest.raises(spk.ReadError):
        spk.CsrMatrix.from_matrix_market(os.fspath(tmp_path / "absent.mtx"))
