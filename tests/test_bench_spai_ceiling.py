"""bench.py's Frobenius-optimal SPAI (the reference point its spai-ceiling
section compares the learned network with): on small Q1 matrices every row
of M satisfies its normal equations, so no change of a row's pattern
entries lowers that row's residual ||e_i - A^T m_i||; and the loss is below
Jacobi's (which is a feasible point of the same minimisation)."""
import importlib.util
import os

import numpy as np
import pytest
import scipy.sparse as sps


@pytest.fixture(scope="module")
def bench():
    p = os.path.join(os.path.dirname(__file__), "..", "bench.py")
    spec = importlib.util.spec_from_file_location("bench_mod", p)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("side,seed", [(12, 3), (20, 4242)])
def test_rows_are_least_squares_optima(O, bench, side, seed):
    R = O.aniso_q1_2d(side, seed, 1.5)
    n = R.nrows
    rp, ci, v = (R.row_ptr.astype(np.int64), R.col_ind.astype(np.int64), R.vals)
    mv = bench.frobenius_optimal_spai(n, rp, ci, v)
    A = sps.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    rng = np.random.default_rng(seed)
    for i in rng.choice(n, 8, replace=False):
        J = ci[rp[i]:rp[i + 1]]
        m = mv[rp[i]:rp[i + 1]]
        e = np.zeros(n)
        e[i] = 1.0
        # the dense least-squares solution over the same pattern
        ml, *_ = np.linalg.lstsq(A[J, :].T, e, rcond=None)
        assert np.allclose(m, ml, rtol=1e-8, atol=1e-10 * np.abs(ml).max())
    M = sps.csr_matrix((mv, ci, rp), shape=(n, n)).toarray()
    d = np.diag(A)
    loss = np.sum((np.eye(n) - M @ A) ** 2) / n
    loss_j = np.sum((np.eye(n) - np.diag(1.0 / d) @ A) ** 2) / n
    assert loss < loss_j
