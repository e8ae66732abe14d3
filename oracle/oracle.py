"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The C sources
restate sparkit's algorithms with the reference's operation order; each
function below names the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPK_ORACLE_LIB: another build of the same sources (the AddressSanitizer +
# UndefinedBehaviorSanitizer one, oracle/_san/liboracle_san.so)
LIB_PATH = os.environ.get("SPK_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

OK, INVALID_ARGUMENT, LOGIC_ERROR, NO_MEMORY, BREAKDOWN = 0, 1, 2, 3, 5


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "-s"],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class CsrD(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64),
                ("row_ptr", C.c_void_p), ("col_ind", C.c_void_p),
                ("vals", C.c_void_p)]


class IluD(C.Structure):
    _fields_ = [("L", CsrD), ("U", CsrD), ("tmp", C.c_void_p)]


class IluJacD(C.Structure):
    _fields_ = [("L", CsrD), ("U", CsrD), ("sweeps", C.c_int64),
                ("tmp", C.c_void_p)]


class ConfigD(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("restart", C.c_int64),
                ("max_iterations", C.c_int64)]


class ConfigF(C.Structure):
    _fields_ = [("tolerance", C.c_float), ("restart", C.c_int64),
                ("max_iterations", C.c_int64)]


class SummaryD(C.Structure):
    _fields_ = [("residual_norm", C.c_double),
                ("computed_iterations", C.c_int64),
                ("converged", C.c_int32), ("status", C.c_int32),
                ("history_len", C.c_int64)]


class SummaryF(C.Structure):
    _fields_ = [("residual_norm", C.c_float),
                ("computed_iterations", C.c_int64),
                ("converged", C.c_int32), ("status", C.c_int32),
                ("history_len", C.c_int64)]


def _declare(L):
    L.orc_csr_pattern.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p,
                                  _i64p, _i64p, _i64p, C.POINTER(C.c_int64)]
    L.orc_csr_pattern.restype = C.c_int
    L.orc_make_matrix.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p,
                                  _i64p, _f64p, _i64p, _i64p, _f64p,
                                  C.POINTER(C.c_int64)]
    L.orc_make_matrix.restype = C.c_int
    L.orc_csr_at.argtypes = [_i64p, _i64p, _f64p, C.c_int64, C.c_int64]
    L.orc_csr_at.restype = C.c_double
    L.orc_spmv_d.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p]
    L.orc_spmv_f.argtypes = [C.c_int64, _i64p, _i64p, _f32p, _f32p, _f32p]
    L.orc_extract_diagonal.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p,
                                       _f64p, _f64p]
    L.orc_jacobi.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, _f64p, _f64p]
    L.orc_jacobi.restype = C.c_int
    L.orc_jacobi_apply_d.argtypes = [C.c_int64, _f64p, _f64p, _f64p]
    L.orc_ilu0.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, _f64p, C.c_int,
                           _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]
    L.orc_ilu0.restype = C.c_int
    for nm, fp in (("d", _f64p), ("f", _f32p)):
        for side in ("forward", "backward"):
            getattr(L, f"orc_{side}_solve_{nm}").argtypes = [
                C.c_int64, _i64p, _i64p, fp, fp, fp]
        getattr(L, f"orc_apply_ilu_jacobi_{nm}").argtypes = [
            C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.orc_true_relres_d.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p,
                                    _f64p]
    L.orc_true_relres_d.restype = C.c_double
    L.orc_splitmix64_at.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_splitmix64_at.restype = C.c_uint64
    L.orc_uniform_pm1.argtypes = [C.c_uint64, C.c_int64, _f64p]
    for nm in ("orc_cg", "orc_gmres", "orc_bicgstab", "orc_cgs", "orc_minres"):
        for sfx, fp, cfg, summ in (("d", _f64p, ConfigD, SummaryD),
                                   ("f", _f32p, ConfigF, SummaryF)):
            getattr(L, f"{nm}_{sfx}").argtypes = [
                C.c_int64, fp, fp, C.POINTER(cfg), C.c_void_p, C.c_void_p,
                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                C.POINTER(summ), C.c_void_p]
    for sfx, fp, cfg, summ in (("d", _f64p, ConfigD, SummaryD),
                               ("f", _f32p, ConfigF, SummaryF)):
        getattr(L, f"orc_fgmres_{sfx}").argtypes = [
            C.c_int64, fp, fp, C.POINTER(cfg), C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_void_p, C.POINTER(summ), C.c_void_p]
    L.orc_gen_tridiagonal.argtypes = [C.c_int64, C.c_double, C.c_double,
                                      C.c_double, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    L.orc_gen_tridiagonal.restype = C.c_int64
    L.orc_gen_poisson_2d.argtypes = [C.c_int64, C.c_int64, C.c_double,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_gen_poisson_2d.restype = C.c_int64
    L.orc_gen_convection_diffusion_2d.argtypes = [
        C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_void_p,
        C.c_void_p, C.c_void_p]
    L.orc_gen_convection_diffusion_2d.restype = C.c_int64
    L.orc_gen_convdiff_centered_2d.argtypes = [
        C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
        C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_gen_convdiff_centered_2d.restype = C.c_int64
    L.orc_gen_nonsymmetric_sample.argtypes = [C.c_void_p] * 3
    L.orc_gen_nonsymmetric_sample.restype = C.c_int64
    L.orc_gen_poisson_3d_csr.argtypes = [C.c_int64] * 3 + [C.c_void_p] * 3
    L.orc_gen_poisson_3d_csr.restype = C.c_int64
    L.orc_gen_convdiff_3d_csr.argtypes = [C.c_int64] * 3 + [C.c_double] * 4 + \
        [C.c_void_p] * 3
    L.orc_gen_convdiff_3d_csr.restype = C.c_int64
    L.orc_powerlaw_cdf.argtypes = [_f64p]
    L.orc_powerlaw_length.argtypes = [_f64p, C.c_uint64, C.c_int64]
    L.orc_powerlaw_length.restype = C.c_int64
    L.orc_gen_powerlaw_spd.argtypes = [C.c_int64, C.c_uint64, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
    L.orc_gen_powerlaw_spd.restype = C.c_int64
    L.orc_aniso_q1_params.argtypes = [C.c_int64, C.c_uint64, C.c_double,
                                      C.c_int, _f64p, _f64p]
    L.orc_gen_aniso_q1_2d.argtypes = [C.c_int64, C.c_uint64, C.c_double,
                                      C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
    L.orc_gen_aniso_q1_2d.restype = C.c_int64
    L.orc_spmv_transpose_d.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p,
                                       _f64p, _f64p, _f64p]
    for sfx, fp, cfg, summ in (("d", _f64p, ConfigD, SummaryD),
                               ("f", _f32p, ConfigF, SummaryF)):
        for nm in ("orc_bicg", "orc_qmr"):
            getattr(L, f"{nm}_{sfx}").argtypes = [
                C.c_int64, fp, fp, C.POINTER(cfg)] + [C.c_void_p] * 8 + [
                C.POINTER(summ), C.c_void_p]
    L.orc_is_valid_permutation.argtypes = [C.c_int64, _i64p]
    L.orc_csr_permute.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p,
                                  C.c_void_p, _i64p, C.c_int, _i64p, _i64p,
                                  C.c_void_p]
    L.orc_batched_side.argtypes = [C.c_uint64, C.c_int64, C.c_int64,
                                   C.c_int64]
    L.orc_batched_side.restype = C.c_int64
    L.orc_batched_kappa.argtypes = [C.c_uint64, C.c_int64, C.c_int64, _f64p]
    L.orc_gen_batched_diffusion.argtypes = [C.c_uint64, C.c_int64, C.c_int64,
                                            C.c_int64, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    L.orc_gen_batched_diffusion.restype = C.c_int64
    L.orc_spai_features.argtypes = [C.POINTER(CsrD), _f32p]
    L.orc_bf16_round.argtypes = [C.c_float]
    L.orc_bf16_round.restype = C.c_float
    L.orc_spai_loss_grad_bf16.argtypes = [_f32p, C.POINTER(CsrD), C.c_void_p]
    L.orc_spai_loss_grad_bf16.restype = C.c_double
    L.orc_spai_loss_grad_exact.argtypes = [_f64p, C.POINTER(CsrD), C.c_void_p]
    L.orc_spai_loss_grad_exact.restype = C.c_double
    L.orc_spai_sym_loss_grad_bf16.argtypes = [_f32p, C.POINTER(CsrD), C.c_void_p]
    L.orc_spai_sym_loss_grad_bf16.restype = C.c_double
    L.orc_spai_sym_loss_grad_exact.argtypes = [_f64p, C.POINTER(CsrD), C.c_void_p]
    L.orc_spai_sym_loss_grad_exact.restype = C.c_double
    L.orc_spai_predict_bf16.argtypes = [_f32p, C.POINTER(CsrD), _f64p]
    L.orc_spai_predict_exact.argtypes = [_f64p, C.POINTER(CsrD), _f64p]


class OracleError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: oracle status {status}")
        self.status = status


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Csr:
    """Host CSR with int64 indices, the reference's layout
    (Compressed_row_sparsity.hpp:97-107, Compressed_row_matrix.hpp:73-76)."""
    nrows: int
    ncols: int
    row_ptr: np.ndarray
    col_ind: np.ndarray
    vals: np.ndarray

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def at(self, r, c):
        return lib().orc_csr_at(self.row_ptr, self.col_ind,
                                np.ascontiguousarray(self.vals, np.float64),
                                r, c)

    def dense(self):
        D = np.zeros((self.nrows, self.ncols))
        for r in range(self.nrows):
            for k in range(self.row_ptr[r], self.row_ptr[r + 1]):
                D[r, self.col_ind[k]] += self.vals[k]
        return D

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.col_ind, self.row_ptr),
                             shape=(self.nrows, self.ncols))

    def _struct(self, vals=None):
        v = self.vals if vals is None else vals
        return CsrD(self.nrows, self.ncols, _ptr(self.row_ptr),
                    _ptr(self.col_ind), _ptr(v))


def csr_pattern(nrows, ncols, rows, cols):
    """Compressed_row_sparsity(Shape, indices):
    Compressed_row_sparsity_Impl.cpp:10-40."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    rp = np.zeros(nrows + 1 if nrows > 0 else 1, np.int64)
    ci = np.zeros(max(len(rows), 1), np.int64)
    nnz = C.c_int64(0)
    st = lib().orc_csr_pattern(nrows, ncols, len(rows), rows, cols, rp, ci,
                               C.byref(nnz))
    if st:
        raise OracleError(st, "csr_pattern")
    return rp, ci[:nnz.value].copy()


def make_matrix(nrows, ncols, rows, cols, vals):
    """make_matrix(shape, entries): matgen.hpp:25-52."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    vals = np.ascontiguousarray(vals, np.float64)
    rp = np.zeros(nrows + 1 if nrows > 0 else 1, np.int64)
    ci = np.zeros(max(len(rows), 1), np.int64)
    v = np.zeros(max(len(rows), 1), np.float64)
    nnz = C.c_int64(0)
    st = lib().orc_make_matrix(nrows, ncols, len(rows), rows, cols, vals, rp,
                               ci, v, C.byref(nnz))
    if st:
        raise OracleError(st, "make_matrix")
    k = nnz.value
    return Csr(nrows, ncols, rp, ci[:k].copy(), v[:k].copy())


def spmv(A: Csr, x):
    """multiply(A, x): sparse_blas.hpp:17-34 (dtype of x selects fp64/fp32)."""
    if np.asarray(x).dtype == np.float32:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(A.nrows, np.float32)
        lib().orc_spmv_f(A.nrows, A.row_ptr, A.col_ind,
                         np.ascontiguousarray(A.vals, np.float32), x, y)
        return y
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(A.nrows, np.float64)
    lib().orc_spmv_d(A.nrows, A.row_ptr, A.col_ind, A.vals, x, y)
    return y


def extract_diagonal(A: Csr):
    """unary.hpp:18-39."""
    d = np.zeros(min(A.nrows, A.ncols), np.float64)
    lib().orc_extract_diagonal(A.nrows, A.ncols, A.row_ptr, A.col_ind, A.vals,
                               d)
    return d


def jacobi(A: Csr):
    """jacobi.hpp:25-36; raises on an exact zero diagonal."""
    d = np.zeros(min(A.nrows, A.ncols), np.float64)
    st = lib().orc_jacobi(A.nrows, A.ncols, A.row_ptr, A.col_ind, A.vals, d)
    if st:
        raise OracleError(st, "jacobi")
    return d


def jacobi_apply(inv_d, r):
    """jacobi.hpp:42-50."""
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros_like(r)
    lib().orc_jacobi_apply_d(len(r), np.ascontiguousarray(inv_d, np.float64),
                             r, z)
    return z


def ilu0(A: Csr, modified: bool = False):
    """ilu0 / milu0: ilu.hpp:35-168.  Returns (L, U)."""
    n = A.nrows
    cap = A.nnz + n + 1
    lrp = np.zeros(n + 1, np.int64)
    urp = np.zeros(n + 1, np.int64)
    lci = np.zeros(cap, np.int64)
    uci = np.zeros(cap, np.int64)
    lv = np.zeros(cap, np.float64)
    uv = np.zeros(cap, np.float64)
    st = lib().orc_ilu0(A.nrows, A.ncols, A.row_ptr, A.col_ind, A.vals,
                        1 if modified else 0, lrp, lci, lv, urp, uci, uv)
    if st:
        raise OracleError(st, "ilu0")
    nl, nu = int(lrp[-1]), int(urp[-1])
    return (Csr(n, n, lrp, lci[:nl].copy(), lv[:nl].copy()),
            Csr(n, n, urp, uci[:nu].copy(), uv[:nu].copy()))


def _ilu_jac_struct(Lf, Uf, k, dtype):
    lv = np.ascontiguousarray(Lf.vals, dtype)
    uv = np.ascontiguousarray(Uf.vals, dtype)
    tmp = np.zeros(3 * Lf.nrows, dtype)
    s = IluJacD(Lf._struct(lv), Uf._struct(uv), int(k), _ptr(tmp))
    return s, [Lf, Uf, lv, uv, tmp, s]


def ilu_jacobi_apply(Lf: Csr, Uf: Csr, k: int, r):
    """Approximate ilu_apply by k Jacobi sweeps per triangle (no reference
    counterpart; replaces the exact solves of ilu.hpp:173-179)."""
    r = np.ascontiguousarray(r)
    dtype = r.dtype
    sfx = "f" if dtype == np.float32 else "d"
    s, keep = _ilu_jac_struct(Lf, Uf, k, dtype)
    out = np.zeros_like(r)
    getattr(lib(), f"orc_apply_ilu_jacobi_{sfx}")(
        C.cast(C.pointer(s), C.c_void_p), len(r), _ptr(r), _ptr(out))
    del keep
    return out


def forward_solve(L: Csr, b):
    """triangular_solve.hpp:27-55."""
    b = np.ascontiguousarray(b)
    x = np.zeros_like(b)
    if b.dtype == np.float32:
        lib().orc_forward_solve_f(L.nrows, L.row_ptr, L.col_ind,
                                  L.vals.astype(np.float32), b, x)
    else:
        lib().orc_forward_solve_d(L.nrows, L.row_ptr, L.col_ind, L.vals,
                                  b.astype(np.float64), x)
    return x


def backward_solve(U: Csr, b):
    """triangular_solve.hpp:67-98."""
    b = np.ascontiguousarray(b)
    x = np.zeros_like(b)
    if b.dtype == np.float32:
        lib().orc_backward_solve_f(U.nrows, U.row_ptr, U.col_ind,
                                   U.vals.astype(np.float32), b, x)
    else:
        lib().orc_backward_solve_d(U.nrows, U.row_ptr, U.col_ind, U.vals,
                                   b.astype(np.float64), x)
    return x


def true_relres(A: Csr, b, x):
    return lib().orc_true_relres_d(A.nrows, A.row_ptr, A.col_ind, A.vals,
                                   np.ascontiguousarray(b, np.float64),
                                   np.ascontiguousarray(x, np.float64))


# ---------------------------------------------------------------- solvers

@dataclass
class Summary:
    """CGSummary / Gmres_summary / Fgmres_summary / Bicgstab_summary
    (conjugate_gradient.hpp:26-40 and siblings) plus a status code."""
    residual_norm: float
    computed_iterations: int
    converged: bool
    status: int
    iteration_residuals: np.ndarray


class _Op:
    """Binds one preconditioner/operator slot to a C callback + context."""

    def __init__(self, spec, A: Csr, dtype):
        L = lib()
        sfx = "f" if dtype == np.float32 else "d"
        self.keep = []
        if spec is None or spec == "identity":
            self.fn = C.cast(getattr(L, f"orc_apply_identity_{sfx}"),
                             C.c_void_p)
            self.ctx = None
            return
        kind, data = spec if isinstance(spec, tuple) else (spec, None)
        if kind in ("csr", "csrT", "csr_ref"):
            M = data if data is not None else A
            v = np.ascontiguousarray(M.vals, dtype)
            s = M._struct(v)
            self.keep += [v, s, M]
            nm = {"csr": "orc_apply_csr", "csr_ref": "orc_apply_csr_ref",
                  "csrT": "orc_apply_csr_transpose"}[kind]
            self.fn = C.cast(getattr(L, f"{nm}_{sfx}"), C.c_void_p)
            self.ctx = C.cast(C.pointer(s), C.c_void_p)
        elif kind == "jacobi":
            inv_d = jacobi(A) if data is None else data
            inv_d = np.ascontiguousarray(inv_d, dtype)
            self.keep.append(inv_d)
            self.fn = C.cast(getattr(L, f"orc_apply_jacobi_{sfx}"),
                             C.c_void_p)
            self.ctx = _ptr(inv_d)
        elif kind in ("ilu", "milu", "ilu_ref"):
            Lf, Uf = data if data is not None else ilu0(A, kind == "milu")
            lv = np.ascontiguousarray(Lf.vals, dtype)
            uv = np.ascontiguousarray(Uf.vals, dtype)
            tmp = np.zeros(A.nrows, dtype)
            s = IluD(Lf._struct(lv), Uf._struct(uv), _ptr(tmp))
            self.keep += [Lf, Uf, lv, uv, tmp, s]
            nm = "orc_apply_ilu_ref" if kind == "ilu_ref" else "orc_apply_ilu"
            self.fn = C.cast(getattr(L, f"{nm}_{sfx}"), C.c_void_p)
            self.ctx = C.cast(C.pointer(s), C.c_void_p)
        elif kind == "ilu_jac":
            # ("ilu_jac", k) or ("ilu_jac", (L, U, k))
            if isinstance(data, tuple):
                Lf, Uf, k = data
            else:
                (Lf, Uf), k = ilu0(A), data
            s, keep = _ilu_jac_struct(Lf, Uf, k, dtype)
            self.keep += keep
            self.fn = C.cast(getattr(L, f"orc_apply_ilu_jacobi_{sfx}"),
                             C.c_void_p)
            self.ctx = C.cast(C.pointer(s), C.c_void_p)
        else:
            raise ValueError(f"unknown operator kind {kind!r}")


def _ref_kind(spec):
    """The reference-call-pattern counterpart of a slot (ILU: a fresh
    vector per triangular solve); other slots allocate nothing in the
    reference and are unchanged."""
    if spec == "ilu":
        return "ilu_ref"
    if isinstance(spec, tuple) and spec[0] == "ilu":
        return ("ilu_ref", spec[1])
    return spec


def _solve(name, A: Csr, b, x0, tol, restart, max_iterations, left, right,
           dtype, ref_calls=False):
    L = lib()
    dtype = np.dtype(dtype)
    sfx = "f" if dtype == np.float32 else "d"
    n = A.nrows
    b = np.ascontiguousarray(b, dtype)
    x = np.zeros(n, dtype) if x0 is None else np.array(x0, dtype=dtype)
    cfg = (ConfigF if sfx == "f" else ConfigD)(tol, restart, max_iterations)
    summ = (SummaryF if sfx == "f" else SummaryD)()
    # two-matvec steps (BiCGSTAB, CGS, BiCG, QMR) can run one matvec past an
    # odd budget, one history entry per matvec
    hist = np.zeros(max(max_iterations, 1) + 2, dtype)
    if ref_calls:  # the reference's per-call allocations and copies
        left, right = _ref_kind(left), _ref_kind(right)
    a_op = _Op(("csr_ref" if ref_calls else "csr", A), A, dtype)
    ml = _Op(left, A, dtype)
    if name in ("bicg", "qmr"):
        at = _Op(("csrT", A), A, dtype)
        mr = _Op(right, A, dtype)
        getattr(L, f"orc_{name}_{sfx}")(n, b, x, C.byref(cfg), a_op.fn, a_op.ctx,
                                      at.fn, at.ctx, ml.fn, ml.ctx, mr.fn,
                                      mr.ctx, C.byref(summ), _ptr(hist))
    elif name == "fgmres":
        getattr(L, f"orc_fgmres_{sfx}")(n, b, x, C.byref(cfg), a_op.fn,
                                        a_op.ctx, ml.fn, ml.ctx,
                                        C.byref(summ), _ptr(hist))
    else:
        mr = _Op(right, A, dtype)
        getattr(L, f"orc_{name}_{sfx}")(n, b, x, C.byref(cfg), a_op.fn,
                                        a_op.ctx, ml.fn, ml.ctx, mr.fn,
                                        mr.ctx, C.byref(summ), _ptr(hist))
    s = Summary(float(summ.residual_norm), int(summ.computed_iterations),
                bool(summ.converged), int(summ.status),
                hist[:summ.history_len].copy())
    return x, s


def cg(A, b, x0=None, *, tol, restart_iterations, max_iterations, left=None,
       right=None, dtype=np.float64, ref_calls=False):
    """conjugate_gradient(): conjugate_gradient.hpp:122-366.  ref_calls:
    the operator slots allocate and copy per call as the reference's
    multiply / ilu_apply do (same results; for timing)."""
    return _solve("cg", A, b, x0, tol, restart_iterations, max_iterations,
                  left, right, dtype, ref_calls)


def gmres(A, b, x0=None, *, tol, restart_dimension, max_iterations, left=None,
          right=None, dtype=np.float64, ref_calls=False):
    """gmres(): gmres.hpp:99-427 (ref_calls as for cg)."""
    return _solve("gmres", A, b, x0, tol, restart_dimension, max_iterations,
                  left, right, dtype, ref_calls)


def fgmres(A, b, x0=None, *, tol, restart_dimension, max_iterations,
           preconditioner=None, dtype=np.float64):
    """fgmres(): fgmres.hpp:96-470."""
    return _solve("fgmres", A, b, x0, tol, restart_dimension, max_iterations,
                  preconditioner, None, dtype)


def bicg(A, b, x0=None, *, tol, max_iterations, left=None, right=None,
         dtype=np.float64):
    """bicg(): bicg.hpp:86-397, transpose operator multiply_transpose."""
    return _solve("bicg", A, b, x0, tol, 0, max_iterations, left, right, dtype)


def minres(A, b, x0=None, *, tol, max_iterations, left=None, right=None,
           dtype=np.float64):
    """minres(): minres.hpp:80-395."""
    return _solve("minres", A, b, x0, tol, 0, max_iterations, left, right,
                  dtype)


def qmr(A, b, x0=None, *, tol, max_iterations, left=None, right=None,
        dtype=np.float64):
    """qmr(): qmr.hpp:86-491, transpose operator multiply_transpose."""
    return _solve("qmr", A, b, x0, tol, 0, max_iterations, left, right, dtype)


def spmv_transpose(A, x):
    """multiply_transpose(A, x): sparse_blas.hpp:36-52."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(A.ncols, np.float64)
    lib().orc_spmv_transpose_d(A.nrows, A.ncols, A.row_ptr, A.col_ind,
                               np.ascontiguousarray(A.vals, np.float64), x, y)
    return y


def cgs(A, b, x0=None, *, tol, max_iterations, left=None, right=None,
        dtype=np.float64):
    """cgs(): cgs.hpp:86-395."""
    return _solve("cgs", A, b, x0, tol, 0, max_iterations, left, right, dtype)


def bicgstab(A, b, x0=None, *, tol, max_iterations, left=None, right=None,
             dtype=np.float64):
    """bicgstab(): bicgstab.hpp:86-410."""
    return _solve("bicgstab", A, b, x0, tol, 0, max_iterations, left, right,
                  dtype)


# ------------------------------------------------------------- generators

def _entries(fn, *args):
    L = lib()
    m = getattr(L, fn)(*args, None, None, None)
    r = np.zeros(m, np.int64)
    c = np.zeros(m, np.int64)
    v = np.zeros(m, np.float64)
    getattr(L, fn)(*args, _ptr(r), _ptr(c), _ptr(v))
    return r, c, v


def tridiagonal_matrix(n, sub, diag, sup):
    """matgen.hpp:104-117."""
    r, c, v = _entries("orc_gen_tridiagonal", n, sub, diag, sup)
    return make_matrix(n, n, r, c, v)


def poisson_2d(nx, ny=None, shift=0.0):
    """poisson_2d(nx, ny[, shift]): matgen.hpp:173-265."""
    ny = nx if ny is None else ny
    r, c, v = _entries("orc_gen_poisson_2d", nx, ny, shift)
    return make_matrix(nx * ny, nx * ny, r, c, v)


def convection_diffusion_2d(nx, ny, d, cx, cy):
    """matgen.hpp:274-325 (upwind)."""
    r, c, v = _entries("orc_gen_convection_diffusion_2d", nx, ny, d, cx, cy)
    return make_matrix(nx * ny, nx * ny, r, c, v)


def convdiff_centered_2d(nx, ny, d, cx, cy, shift=0.0):
    """matgen.hpp:339-386."""
    r, c, v = _entries("orc_gen_convdiff_centered_2d", nx, ny, d, cx, cy,
                       shift)
    return make_matrix(nx * ny, nx * ny, r, c, v)


def nonsymmetric_sample():
    """matgen.hpp:147-164."""
    r, c, v = _entries("orc_gen_nonsymmetric_sample")
    return make_matrix(4, 4, r, c, v)


def _csr_gen(fn, n, *args):
    L = lib()
    nnz = getattr(L, fn)(*args, None, None, None)
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(nnz, np.int64)
    v = np.zeros(nnz, np.float64)
    getattr(L, fn)(*args, _ptr(rp), _ptr(ci), _ptr(v))
    return Csr(n, n, rp, ci, v)


def poisson_3d(nx, ny=None, nz=None):
    """7-point Dirichlet Laplacian (no reference counterpart)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _csr_gen("orc_gen_poisson_3d_csr", nx * ny * nz, nx, ny, nz)


def convdiff_3d(nx, ny, nz, d, cx, cy, cz):
    """7-point upwind convection-diffusion, Dirichlet diagonal (no reference
    counterpart; extends matgen.hpp:292-321 to 3D)."""
    return _csr_gen("orc_gen_convdiff_3d_csr", nx * ny * nz, nx, ny, nz, d,
                    cx, cy, cz)


def powerlaw_cdf():
    cdf = np.zeros(4096 - 3 + 1, np.float64)
    lib().orc_powerlaw_cdf(cdf)
    return cdf


def powerlaw_lengths(n, seed):
    cdf = powerlaw_cdf()
    return np.array([lib().orc_powerlaw_length(cdf, seed, i)
                     for i in range(n)], np.int64)


def powerlaw_spd(n, seed):
    """Power-law SPD matrix (no reference counterpart; definition in
    oracle_gen.c)."""
    L = lib()
    nnz = L.orc_gen_powerlaw_spd(n, seed, None, None, None)
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(nnz, np.int64)
    v = np.zeros(nnz, np.float64)
    L.orc_gen_powerlaw_spd(n, seed, _ptr(rp), _ptr(ci), _ptr(v))
    return Csr(n, n, rp, ci, v)


def aniso_q1_params(N, seed, sigma=1.5, isotropic=False):
    """Cell conductivities ((N+1)^2) and the 4x4 element table of
    aniso_q1_2d."""
    kappa = np.zeros((N + 1) * (N + 1), np.float64)
    E = np.zeros(16, np.float64)
    lib().orc_aniso_q1_params(N, seed, sigma, int(isotropic), kappa, E)
    return kappa, E.reshape(4, 4)


def aniso_q1_2d(N, seed, sigma=1.5, isotropic=False):
    """2D anisotropic variable-coefficient diffusion, Q1 finite elements on
    N x N interior nodes (no reference counterpart; definition in
    oracle_gen.c)."""
    n = N * N
    nnz = (3 * N - 2) ** 2
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(nnz, np.int64)
    v = np.zeros(nnz, np.float64)
    got = lib().orc_gen_aniso_q1_2d(N, seed, sigma, int(isotropic), _ptr(rp),
                                    _ptr(ci), _ptr(v))
    assert got == nnz
    return Csr(n, n, rp, ci, v)


def symmetric_part(M: Csr) -> Csr:
    """(M + M^T) / 2 on M's pattern, which must be structurally symmetric
    (no reference counterpart)."""
    n = M.nrows
    pos = {}
    for i in range(n):
        for k in range(M.row_ptr[i], M.row_ptr[i + 1]):
            pos[(i, int(M.col_ind[k]))] = k
    v = np.empty(M.nnz)
    for (i, j), k in pos.items():
        if (j, i) not in pos:
            raise OracleError(2, "pattern is not structurally symmetric")
        v[k] = 0.5 * (M.vals[k] + M.vals[pos[(j, i)]])
    return Csr(n, M.ncols, M.row_ptr.copy(), M.col_ind.copy(), v)


def _mm_int(tok):
    """`is >> size_type` on one token: the leading integer, else 0."""
    import re
    m = re.match(r"[+-]?\d+", tok)
    return int(m.group(0)) if m else 0


def read_matrix_market_text(text):
    """read_matrix_market (matrix_market.hpp:43-104) + to_compressed_row:
    (rows, cols, {(i, j): value}) with the last entry at a position kept
    (Coordinate_matrix::insert_or_assign), as a Csr."""
    lines = text.split("\n")
    if text.endswith("\n"):
        lines = lines[:-1]
    it = iter(lines)
    try:
        banner = next(it)
    except StopIteration:
        raise OracleError(7, "matrix market: unexpected end of input")
    tok = (banner.split() + [""] * 5)[:5]
    if tok[0].lower() != "%%matrixmarket":
        raise OracleError(7, "invalid banner header")
    fmt, field, sym = tok[2].lower(), tok[3].lower(), tok[4].lower()
    if fmt not in ("coordinate", "array") or field not in (
            "real", "integer", "complex", "pattern") or sym not in (
            "general", "symmetric", "skew-symmetric", "hermitian"):
        raise OracleError(7, "unknown banner word")
    if fmt != "coordinate" or field == "complex":
        raise OracleError(7, "unsupported")
    line = ""
    for line in it:
        if line and line[0] != "%":
            break
    t = line.split() + ["", "", ""]
    rows, cols, nnz = _mm_int(t[0]), _mm_int(t[1]), _mm_int(t[2])
    entries = {}
    for _ in range(nnz):
        e = (next(it, "").split() + ["", "", ""])
        i, j = _mm_int(e[0]) - 1, _mm_int(e[1]) - 1
        if field == "pattern":
            v = 1.0
        else:
            try:
                v = float(e[2]) if e[2][:1] in "+-.0123456789" and e[2] else 0.0
            except ValueError:
                v = 0.0
        entries[(i, j)] = v
        if sym == "symmetric" and i != j:
            entries[(j, i)] = v
    keys = list(entries)
    return make_matrix(rows, cols, [k[0] for k in keys], [k[1] for k in keys],
                       [entries[k] for k in keys])


def write_matrix_market_text(A: Csr):
    """write_matrix_market(os, A) (matrix_market.hpp:123-148): values as
    std::ostream prints a double by default (%g)."""
    out = ["%%MatrixMarket matrix coordinate real general",
           f"{A.nrows} {A.ncols} {A.nnz}"]
    for i in range(A.nrows):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            out.append(f"{i + 1} {A.col_ind[k] + 1} {'%g' % A.vals[k]}")
    return "\n".join(out) + "\n"


def symmetrize_pattern(A: Csr):
    """symmetrize_pattern (reordering.cpp:19-38): the pattern of A + A^T
    (sorted unique columns per row), returned as (row_ptr, col_ind)."""
    n = A.nrows
    rows = [set() for _ in range(n)]
    for i in range(n):
        for k in range(A.row_ptr[i], A.row_ptr[i + 1]):
            j = int(A.col_ind[k])
            rows[i].add(j)
            rows[j].add(i)
    rp = np.zeros(n + 1, np.int64)
    ci = []
    for i in range(n):
        r = sorted(rows[i])
        ci.extend(r)
        rp[i + 1] = rp[i] + len(r)
    return rp, np.array(ci, np.int64)


def adjacency_degree(rp, ci):
    """adjacency_degree (reordering.cpp:40-56): off-diagonal count per row."""
    n = len(rp) - 1
    return np.array([sum(1 for k in range(rp[i], rp[i + 1]) if ci[k] != i)
                     for i in range(n)], np.int64)


def _cm_bfs(rp, ci, deg, start, visited):
    """One Cuthill-McKee sweep from start over unvisited nodes: each node's
    unvisited neighbours are appended sorted by degree, ties in column order
    (reordering.cpp:62-108, :169-191).  Returns the order and the level of
    every appended node."""
    order, level = [start], [0]
    visited[start] = True
    head = 0
    while head < len(order):
        u = order[head]
        nb = [int(ci[k]) for k in range(rp[u], rp[u + 1])
              if ci[k] != u and not visited[ci[k]]]
        nb.sort(key=lambda v: deg[v])  # stable
        for v in nb:
            if not visited[v]:
                visited[v] = True
                order.append(v)
                level.append(level[head] + 1)
        head += 1
    return order, level


def pseudo_peripheral_node(A: Csr):
    """pseudo_peripheral_node (reordering.cpp:112-150)."""
    rp, ci = symmetrize_pattern(A)
    deg = adjacency_degree(rp, ci)
    start = int(np.argmin(deg))  # the first node of minimum degree

    def levels(s):
        order, lev = _cm_bfs(rp, ci, deg, s, np.zeros(A.nrows, bool))
        ecc = max(lev)
        return ecc, [v for v, l in zip(order, lev) if l == ecc]

    ecc, last = levels(start)
    while True:
        cand = last[0]
        for v in last[1:]:
            if deg[v] < deg[cand]:
                cand = v
        new_ecc, new_last = levels(cand)
        if new_ecc <= ecc:
            return start
        start, ecc, last = cand, new_ecc, new_last


def reverse_cuthill_mckee(A: Csr):
    """reverse_cuthill_mckee (reordering.cpp:152-203): perm[old] = new."""
    n = A.nrows
    rp, ci = symmetrize_pattern(A)
    deg = adjacency_degree(rp, ci)
    visited = np.zeros(n, bool)
    ordering = []
    while len(ordering) < n:
        if not ordering:
            start = pseudo_peripheral_node(A)
        else:
            cand = [i for i in range(n) if not visited[i]]
            start = min(cand, key=lambda i: (deg[i], i))
        order, _ = _cm_bfs(rp, ci, deg, start, visited)
        ordering.extend(order)
    ordering.reverse()
    perm = np.zeros(n, np.int64)
    for new, old in enumerate(ordering):
        perm[old] = new
    return perm


def is_valid_permutation(perm):
    return bool(lib().orc_is_valid_permutation(
        len(perm), np.ascontiguousarray(perm, np.int64)))


def permute(A, perm, mode):
    """rperm (mode 'rows'), cperm ('cols') or dperm ('both'):
    permutation.hpp:41-148 (perm[old] = new)."""
    m = {"rows": 1, "cols": 2, "both": 3}[mode]
    perm = np.ascontiguousarray(perm, np.int64)
    rp = np.zeros(A.nrows + 1, np.int64)
    ci = np.zeros(max(A.nnz, 1), np.int64)
    v = np.zeros(max(A.nnz, 1), np.float64)
    vals = np.ascontiguousarray(A.vals, np.float64)
    st = lib().orc_csr_permute(A.nrows, A.ncols, A.row_ptr, A.col_ind,
                               _ptr(vals), perm, m, rp, ci, _ptr(v))
    if st:
        raise OracleError(st, "permute")
    return Csr(A.nrows, A.ncols, rp, ci[:A.nnz].copy(), v[:A.nnz].copy())


def batched_side(seed, k, smin=32, smax=256):
    return int(lib().orc_batched_side(seed, k, smin, smax))


def batched_kappa(seed, k, smin=32, smax=256):
    s = batched_side(seed, k, smin, smax)
    kap = np.zeros(s * s, np.float64)
    lib().orc_batched_kappa(seed, k, s, kap)
    return kap.reshape(s, s)


def batched_diffusion(seed, k, smin=32, smax=256):
    """System k of the batched variable-coefficient diffusion family (no
    reference counterpart; definition in oracle_gen.c)."""
    s = batched_side(seed, k, smin, smax)
    n = s * s
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(5 * n, np.int64)
    v = np.zeros(5 * n, np.float64)
    nnz = lib().orc_gen_batched_diffusion(seed, k, smin, smax, _ptr(rp),
                                          _ptr(ci), _ptr(v))
    return Csr(n, n, rp, ci[:nnz].copy(), v[:nnz].copy())


# ---------------------------------------------------------------- learned SPAI
SPAI_NPARAM = 6345
SPAI_SHAPES = [("W1", (24, 64)), ("b1", (64,)), ("W2", (64, 64)),
               ("b2", (64,)), ("W3", (64, 9)), ("b3", (9,))]


def spai_unpack(params):
    out, k = {}, 0
    for name, shp in SPAI_SHAPES:
        m = int(np.prod(shp))
        out[name] = params[k:k + m].reshape(shp)
        k += m
    return out


def spai_features(A):
    X = np.zeros((A.nrows, 24), np.float32)
    st = lib().orc_spai_features(C.byref(A._struct()), X)
    if st:
        raise OracleError(st, "spai_features")
    return X


def spai_loss_grad(params, A, exact=False, symmetric=False):
    """(loss, gradient) of (1/n)||I - M A||_F^2, or with symmetric of
    (1/n)||I - S A||_F^2 for S = (M + M^T) / 2 (no reference counterpart:
    the objective of the operator CG applies); bf16 device arithmetic
    unless exact (float64 throughout)."""
    g = np.zeros(SPAI_NPARAM, np.float64)
    name = "orc_spai_sym_loss_grad" if symmetric else "orc_spai_loss_grad"
    if exact:
        P = np.ascontiguousarray(params, np.float64)
        loss = getattr(lib(), name + "_exact")(P, C.byref(A._struct()), _ptr(g))
    else:
        P = np.ascontiguousarray(params, np.float32)
        loss = getattr(lib(), name + "_bf16")(P, C.byref(A._struct()), _ptr(g))
    return loss, g


def spai_predict(params, A, exact=False):
    """M's values on A's pattern."""
    m = np.zeros(A.nnz, np.float64)
    if exact:
        st = lib().orc_spai_predict_exact(
            np.ascontiguousarray(params, np.float64), C.byref(A._struct()), m)
    else:
        st = lib().orc_spai_predict_bf16(
            np.ascontiguousarray(params, np.float32), C.byref(A._struct()), m)
    if st:
        raise OracleError(st, "spai_predict")
    return m


def bf16_round(x):
    return np.array([lib().orc_bf16_round(float(v)) for v in np.ravel(x)],
                    np.float32).reshape(np.shape(x))


def splitmix64_at(seed, k):
    return int(lib().orc_splitmix64_at(seed, k))


def uniform_pm1(seed, n):
    out = np.zeros(n, np.float64)
    lib().orc_uniform_pm1(seed, n, out)
    return out
