"""Pins the CPU oracle's CSR construction, SpMV, Jacobi, ILU(0), triangular
solves and generators against the reference's own known-answer tests
(sparkit_testing/*.cpp), closed forms, brute force and scipy."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla


# --------------------------------------------------------------- pattern

def test_pattern_kats(O):
    # compressed_data_test.cpp:142-154: columns sorted within rows
    rp, ci = O.csr_pattern(4, 6, [2, 2, 2], [5, 2, 4])
    assert ci.tolist() == [2, 4, 5]
    # :156-167: duplicates collapsed
    rp, ci = O.csr_pattern(4, 5, [2, 2, 3], [3, 3, 4])
    assert len(ci) == 2 and rp[3] == 1 and rp[4] == 2
    # :169-181: empty matrix
    rp, ci = O.csr_pattern(3, 3, [], [])
    assert rp.tolist() == [0, 0, 0, 0] and len(ci) == 0
    # :183-203: single row, many entries
    rp, ci = O.csr_pattern(4, 8, [2, 2, 2, 2], [7, 3, 2, 5])
    assert rp.tolist() == [0, 0, 0, 4, 4] and ci.tolist() == [2, 3, 5, 7]


def test_pattern_errors(O):
    with pytest.raises(O.OracleError) as e:   # Index.cpp:8
        O.csr_pattern(3, 3, [0, -1], [0, 1])
    assert e.value.status == O.LOGIC_ERROR
    with pytest.raises(O.OracleError) as e:   # Shape.cpp:8-10
        O.csr_pattern(1, 3, [0], [0])
    assert e.value.status == O.LOGIC_ERROR


@pytest.mark.parametrize("seed", range(5))
def test_pattern_brute_force(O, seed):
    """Sorted unique (row, col) pairs: a python set + sort is the
    definition of Compressed_row_sparsity_Impl.cpp:10-40."""
    rng = np.random.default_rng(seed)
    nr, nc = int(rng.integers(2, 40)), int(rng.integers(2, 40))
    m = int(rng.integers(0, 400))
    rows, cols = rng.integers(0, nr, m), rng.integers(0, nc, m)
    rp, ci = O.csr_pattern(nr, nc, rows, cols)
    pairs = sorted(set(zip(rows.tolist(), cols.tolist())))
    assert ci.tolist() == [c for _, c in pairs]
    counts = np.bincount([r for r, _ in pairs], minlength=nr)
    assert rp.tolist() == [0] + np.cumsum(counts).tolist()


def test_make_matrix_round_trip_and_last_wins(O):
    # matgen_test.cpp:49-72
    A = O.make_matrix(3, 3, [0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [1., 2., 3., 4., 5.])
    assert A.nnz == 5
    assert [A.at(0, 0), A.at(0, 2), A.at(1, 1), A.at(2, 0), A.at(2, 2)] == \
        [1., 2., 3., 4., 5.]
    assert A.at(0, 1) == 0.0 and A.at(1, 0) == 0.0
    # matgen.hpp:40-49: a later duplicate overwrites
    B = O.make_matrix(2, 2, [0, 0, 1], [1, 1, 0], [7., 9., 3.])
    assert B.at(0, 1) == 9.0 and B.nnz == 2


# ------------------------------------------------------------ generators

@pytest.mark.parametrize("N", [2, 3, 7, 16])
def test_stencil_nnz_closed_forms(O, N):
    # 5-point: 5N^2 - 4N; 7-point: 7N^3 - 6N^2
    assert O.poisson_2d(N).nnz == 5 * N * N - 4 * N
    assert O.poisson_3d(N).nnz == 7 * N ** 3 - 6 * N * N


def test_poisson_2d_kats(O):
    # matgen_test.cpp:154-194
    A = O.poisson_2d(4)
    assert A.nrows == 16
    assert (A.at(0, 0), A.at(0, 1), A.at(0, 4)) == (2.0, -1.0, -1.0)
    i = 5
    assert A.at(i, i) == 4.0
    assert all(A.at(i, j) == -1.0 for j in (i - 1, i + 1, i - 4, i + 4))
    assert A.at(1, 1) == 3.0
    R = O.poisson_2d(3, 5)
    assert (R.at(0, 0), R.at(1, 1), R.at(4, 4)) == (2.0, 3.0, 4.0)
    # singular Neumann Laplacian: zero row sums (matgen.hpp:173-208)
    assert np.all(O.spmv(O.poisson_2d(9), np.ones(81)) == 0.0)


def test_shifted_and_convdiff_kats(O):
    # matgen_test.cpp:348-379 (make_grid_16)
    A = O.poisson_2d(4, 4, 5.0)
    assert (A.at(0, 0), A.at(1, 1), A.at(5, 5)) == (7.0, 8.0, 9.0)
    S = A.to_scipy().toarray()
    assert np.array_equal(S, S.T)
    # matgen_test.cpp:385-414
    C = O.convdiff_centered_2d(4, 4, 1.0, 0.3, 0.3, 5.0)
    n = 5
    assert C.at(n, n - 1) == pytest.approx(-1.3)
    assert C.at(n, n + 1) == pytest.approx(-0.7)
    assert C.at(n, n - 4) == pytest.approx(-1.3)
    assert C.at(n, n + 4) == pytest.approx(-0.7)
    assert C.at(n, n) == 9.0
    # :416-436 zero convection reduces to shifted poisson
    Z = O.convdiff_centered_2d(4, 4, 1.0, 0.0, 0.0, 5.0)
    assert np.array_equal(Z.to_scipy().toarray(), A.to_scipy().toarray())
    # matgen_test.cpp:243-262: upwind with zero convection == poisson_2d
    U = O.convection_diffusion_2d(5, 4, 1.0, 0.0, 0.0)
    assert np.array_equal(U.to_scipy().toarray(),
                          O.poisson_2d(5, 4).to_scipy().toarray())
    # upwind: zero row sums (singular), nonsymmetric
    W = O.convection_diffusion_2d(6, 5, 1.0, 0.5, 0.25)
    assert np.allclose(O.spmv(W, np.ones(30)), 0.0, atol=1e-14)


def test_nonsymmetric_sample_kats(O):
    # matgen_test.cpp:438-471
    A = O.nonsymmetric_sample()
    expect = {(0, 0): 4, (0, 1): -1, (0, 2): .5, (1, 0): -.5, (1, 1): 4,
              (1, 3): -1, (2, 0): .3, (2, 2): 4, (2, 3): -.8, (3, 1): -.6,
              (3, 2): .2, (3, 3): 4}
    for (i, j) in itertools.product(range(4), range(4)):
        assert A.at(i, j) == expect.get((i, j), 0.0)


def test_3d_generators_closed_form(O):
    A = O.poisson_3d(5, 4, 3)
    D = A.to_scipy().toarray()
    assert np.all(np.diag(D) == 6.0) and np.array_equal(D, D.T)
    # interior rows sum to zero, boundary rows are strictly dominant
    C = O.convdiff_3d(5, 4, 3, 1.0, 0.3, 0.2, 0.1)
    M = C.to_scipy().toarray()
    assert np.all(np.diag(M) == (6.0 + 0.3) + 0.2 + 0.1)
    row = 1 * 20 + 1 * 5 + 2  # interior point (i=2, j=1, k=1)
    assert M[row, row - 1] == -1.3 and M[row, row + 1] == -1.0
    assert M[row, row - 5] == -1.2 and M[row, row - 20] == -1.1
    assert abs(M[row].sum()) < 1e-14


def test_splitmix64_known_values(O):
    # splitmix64 with state 0: first outputs of the published generator
    assert O.splitmix64_at(0, 0) == 0xE220A8397B1DCDAF
    assert O.splitmix64_at(0, 1) == 0x6E789E6AA1B965F4
    u = O.uniform_pm1(42, 100000)
    assert -1.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean()) < 0.01


# ------------------------------------------------------------------ SpMV

def test_spmv_kats(O):
    # sparse_blas_test.cpp:36-95
    A = O.make_matrix(3, 3, [0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [2., 3., 4., 5., 6.])
    assert O.spmv(A, np.array([1., 2., 3.])).tolist() == [11., 8., 23.]
    E = O.make_matrix(3, 3, [], [], [])
    assert O.spmv(E, np.array([1., 2., 3.])).tolist() == [0., 0., 0.]
    R = O.make_matrix(2, 3, [0, 0, 1], [0, 2, 1], [1., 2., 3.])
    assert O.spmv(R, np.array([1., 2., 3.])).tolist() == [7., 6.]


@pytest.mark.parametrize("seed", range(3))
def test_spmv_exact_integers_and_scipy(O, seed):
    rng = np.random.default_rng(seed)
    n = 300
    rows = rng.integers(0, n, 3000)
    cols = rng.integers(0, 250, 3000)
    vals = rng.integers(-9, 10, 3000).astype(float)
    A = O.make_matrix(n, 250, rows, cols, vals)
    xi = rng.integers(-9, 10, 250).astype(float)
    # integer data: every partial sum is exact, so the result is exact
    assert np.array_equal(O.spmv(A, xi), A.to_scipy() @ xi)
    xf = rng.uniform(-1, 1, 250)
    assert np.allclose(O.spmv(A, xf), A.to_scipy() @ xf, rtol=1e-13,
                       atol=1e-13)
    # sequential order: identical to an explicit python left-to-right sum
    r = int(np.argmax(np.diff(A.row_ptr)))
    s = 0.0
    for p in range(A.row_ptr[r], A.row_ptr[r + 1]):
        s += A.vals[p] * xf[A.col_ind[p]]
    assert O.spmv(A, xf)[r] == s


# ----------------------------------------------------- diagonal / Jacobi

def test_extract_diagonal_kats(O):
    # unary_test.cpp:35-95
    A = O.make_matrix(3, 3, [0, 0, 1, 1, 2, 2], [0, 1, 1, 2, 0, 2],
                      [1., 2., 3., 4., 5., 6.])
    assert O.extract_diagonal(A).tolist() == [1., 3., 6.]
    R = O.make_matrix(2, 4, [0, 0, 1, 1], [0, 3, 1, 2], [7., 8., 9., 10.])
    assert O.extract_diagonal(R).tolist() == [7., 9.]
    E = O.make_matrix(3, 3, [], [], [])
    assert O.extract_diagonal(E).tolist() == [0., 0., 0.]
    Z = O.make_matrix(2, 2, [0, 1], [1, 0], [5., 6.])
    assert O.extract_diagonal(Z).tolist() == [0., 0.]


def test_jacobi_kats(O):
    # jacobi_test.cpp:50-95
    D = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    assert O.jacobi(D).tolist() == [0.5, 1 / 3, 0.25, 0.2]
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    assert O.jacobi(T).tolist() == [0.25] * 4
    D2 = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 4., 5., 10.])
    z = O.jacobi_apply(O.jacobi(D2), np.array([6., 12., 20., 30.]))
    assert z.tolist() == [3., 3., 4., 3.]
    # jacobi_test.cpp:273-282
    Z = O.make_matrix(3, 3, [0, 0, 1, 2], [0, 1, 0, 2], [2., 1., 1., 3.])
    with pytest.raises(O.OracleError) as e:
        O.jacobi(Z)
    assert e.value.status == O.INVALID_ARGUMENT


# --------------------------------------------------------------- ILU(0)

def _dense_ilu0(A, modified=False):
    """IKJ ILU(0) on A's pattern over a dense array (independent of the
    CSR bookkeeping): returns the factored dense array."""
    W = A.to_scipy().toarray()
    mask = W != 0
    mask |= A.to_scipy().astype(bool).toarray()
    n = W.shape[0]
    for i in range(n):
        dropped = 0.0
        for k in range(i):
            if not mask[i, k]:
                continue
            W[i, k] = W[i, k] / W[k, k]
            l = W[i, k]
            for j in range(k + 1, n):
                if mask[k, j]:
                    if mask[i, j]:
                        W[i, j] = W[i, j] - l * W[k, j]
                    elif modified:
                        dropped = dropped + l * W[k, j]
        if modified:
            W[i, i] = W[i, i] - dropped
    return W, mask


def _factor_cases(O):
    return [
        O.tridiagonal_matrix(8, -1.0, 4.0, -1.0),
        O.nonsymmetric_sample(),
        O.convdiff_centered_2d(5, 4, 1.0, 0.3, 0.2, 5.0),
        O.convection_diffusion_2d(6, 5, 1.0, 0.45, 0.2),
        O.poisson_2d(5, 5, 0.5),
        O.convdiff_3d(4, 3, 3, 1.0, 0.4, 0.2, 0.1),
    ]


@pytest.mark.parametrize("modified", [False, True])
def test_ilu0_matches_dense_ikj(O, modified):
    for A in _factor_cases(O):
        L, U = O.ilu0(A, modified)
        W, mask = _dense_ilu0(A, modified)
        Ld = L.to_scipy().toarray()
        Ud = U.to_scipy().toarray()
        n = A.nrows
        assert np.array_equal(np.tril(Ld, -1), np.tril(W, -1) * np.tril(mask, -1))
        assert np.array_equal(np.diag(Ld), np.ones(n))
        assert np.array_equal(Ud, np.triu(W) * np.triu(mask))
        # layout: L's diagonal is the last entry of each row, U's the first
        assert all(L.col_ind[L.row_ptr[i + 1] - 1] == i for i in range(n))
        assert all(U.col_ind[U.row_ptr[i]] == i for i in range(n))


def test_ilu0_defining_property(O):
    """ILU(0): (L U)_ij = a_ij on A's pattern; MILU(0): (L U) e = A e."""
    for A in _factor_cases(O):
        S = A.to_scipy()
        L, U = O.ilu0(A)
        LU = L.to_scipy() @ U.to_scipy()
        r, c = S.nonzero()
        assert np.allclose(np.asarray(LU[r, c]).ravel(),
                           np.asarray(S[r, c]).ravel(), rtol=1e-13, atol=1e-13)
        Lm, Um = O.ilu0(A, True)
        e = np.ones(A.nrows)
        assert np.allclose(Lm.to_scipy() @ (Um.to_scipy() @ e), S @ e,
                           rtol=1e-12, atol=1e-12)


def test_ilu0_kats(O):
    # ilu_test.cpp:54-74: diagonal -> L = I, U = diag(A)
    D = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    L, U = O.ilu0(D)
    assert L.vals.tolist() == [1.] * 4 and U.vals.tolist() == [2., 3., 4., 5.]
    # :76-107 tridiag: exact LU, so ilu_apply solves exactly
    T = O.tridiagonal_matrix(4, -1.0, 4.0, -1.0)
    L, U = O.ilu0(T)
    assert np.allclose((L.to_scipy() @ U.to_scipy()).toarray(),
                       T.to_scipy().toarray(), atol=1e-12)
    b = O.spmv(T, np.array([1., 2., 3., 4.]))
    z = O.backward_solve(U, O.forward_solve(L, b))
    assert np.allclose(z, [1., 2., 3., 4.], atol=1e-12)
    # :288-307: milu == ilu when there is no fill
    Lm, Um = O.ilu0(T, True)
    assert np.array_equal(Lm.vals, L.vals) and np.array_equal(Um.vals, U.vals)
    # :274-282 non-square; ilu.hpp:63-65 missing diagonal
    R = O.make_matrix(3, 4, [0, 1, 2], [0, 1, 2], [1., 2., 3.])
    with pytest.raises(O.OracleError):
        O.ilu0(R)
    M = O.make_matrix(3, 3, [0, 1, 1, 2], [0, 0, 2, 2], [1., 1., 1., 1.])
    with pytest.raises(O.OracleError):
        O.ilu0(M)


# ---------------------------------------------------- triangular solves

def test_triangular_solve_kats(O):
    # triangular_solve_test.cpp:72-126, 219-240
    I = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [1.] * 4)
    assert O.forward_solve(I, np.array([1., 2., 3., 4.])).tolist() == [1., 2., 3., 4.]
    D = O.make_matrix(4, 4, [0, 1, 2, 3], [0, 1, 2, 3], [2., 3., 4., 5.])
    assert O.forward_solve(D, np.array([6., 12., 20., 30.])).tolist() == [3., 4., 5., 6.]
    L2 = O.make_matrix(2, 2, [0, 1, 1], [0, 0, 1], [2., 1., 2.])
    assert O.forward_solve(L2, np.array([4., 5.])).tolist() == [2.0, 1.5]
    U2 = O.make_matrix(2, 2, [0, 0, 1], [0, 1, 1], [2., 1., 2.])
    assert O.backward_solve(U2, np.array([5., 4.])).tolist() == [1.5, 2.0]


def test_triangular_solves_vs_dense(O):
    for A in _factor_cases(O):
        L, U = O.ilu0(A)
        b = O.uniform_pm1(5, A.nrows)
        y = O.forward_solve(L, b)
        assert np.allclose(y, sla.solve_triangular(L.to_scipy().toarray(), b,
                                                   lower=True), rtol=1e-12, atol=1e-12)
        z = O.backward_solve(U, b)
        assert np.allclose(z, sla.solve_triangular(U.to_scipy().toarray(), b,
                                                   lower=False), rtol=1e-11, atol=1e-11)
