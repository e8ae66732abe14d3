"""sparkit_amd -- MI355X (gfx950) implementation of sparkit's CSR SpMV +
preconditioned Krylov solve path.

Python mirror of the reference's C++ interface (namespace
sparkit::data::detail), over the C ABI in include/sparkit_hip.h:

  CsrMatrix              Compressed_row_matrix<T> held in HBM
                         (Compressed_row_matrix.hpp:20-132)
  multiply(A, x)         sparse_blas.hpp:17-34
  extract_diagonal(A)    unary.hpp:18-39
  jacobi(A), jacobi_apply                jacobi.hpp:25-50
  ilu0(A), milu0(A), ilu_apply           ilu.hpp:35-179
  conjugate_gradient(b, x, cfg, A, M_L, M_R)   conjugate_gradient.hpp:339-366
  gmres(b, x, cfg, A, M_L, M_R)                gmres.hpp:400-427
  fgmres(b, x, cfg, A, M)                      fgmres.hpp:452-470
  bicgstab(b, x, cfg, A, M_L, M_R)             bicgstab.hpp:383-410

Operators and preconditioners are device objects; vectors are
DeviceVector (HBM).  Everything runs on the HIP library; there is no CPU
path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (F32, F64, BREAKDOWN, ORTHO_CGS1, ORTHO_CGS2, ORTHO_MGS,
                   SPAI_LOSS_INVERSE,
                   SPAI_LOSS_SYMMETRIC, InvalidArgument,
                   LogicError, ReadError, SpkError, call, load)

__all__ = [
    "init", "synchronize", "DeviceVector", "CsrMatrix", "Preconditioner",
    "identity", "jacobi", "ilu0", "milu0", "spai", "multiply",
    "extract_diagonal", "jacobi_apply", "ilu_apply", "CGConfig",
    "GmresConfig", "FgmresConfig", "BicgstabConfig", "Summary",
    "conjugate_gradient", "conjugate_gradient_noalloc", "gmres", "fgmres", "bicgstab", "cgs", "CgsConfig",
    "bicg", "BicgConfig", "minres", "MinresConfig", "qmr", "QmrConfig",
    "true_relres", "dot", "axpby", "inner_cg", "inner_ir", "InnerCg", "learned_solve", "LearnedConfig",
    "LearnedSummary", "FALLBACK_AUTO", "FALLBACK_JACOBI_PCG",
    "FALLBACK_ILU_GMRES", "FALLBACK_NONE",
    "F32", "F64", "ORTHO_MGS", "ORTHO_CGS2", "ORTHO_CGS1", "SPAI_LOSS_INVERSE", "SPAI_LOSS_SYMMETRIC", "SpkError", "InvalidArgument", "LogicError", "ReadError", "Timer", "gather_rate",
    "IrConfig", "IrSummary", "ir_solve", "IR_GMRES", "IR_FGMRES", "IR_CG",
    "IR_BICGSTAB", "IR_CG_RELIABLE",
    "SpaiModel", "spai_features", "SPAI_NPARAM", "BatchConfig",
    "BatchedSystems",
    "poisson_2d", "poisson_3d", "convdiff_3d", "powerlaw_spd", "aniso_q1_2d",
    "uniform_pm1",
]

_DT = {np.dtype(np.float64): F64, np.dtype(np.float32): F32}


def _code(dtype) -> int:
    return _DT[np.dtype(dtype)]


def init(device: int = 0):
    call("spk_init", device)


def synchronize():
    call("spk_synchronize")


def device_info():
    cus = C.c_int(0)
    buf = C.create_string_buffer(256)
    call("spk_device_info", C.byref(cus), buf, 256)
    return cus.value, buf.value.decode()


def _gfx_target():
    """The GPU target name (e.g. gfx950) from spk_device_info."""
    name = device_info()[1]
    for tok in name.replace("(", " ").replace(")", " ").replace(":", " ").split():
        if tok.startswith("gfx"):
            return tok
    return name


class Timer:
    """HIP events on the library stream (the stream every kernel uses)."""

    def start(self):
        call("spk_timer_start")

    def stop(self) -> float:
        ms = C.c_float(0)
        call("spk_timer_stop", C.byref(ms))
        return ms.value


def gather_rate(window: int, count: int = 64 << 20, reps: int = 3):
    """Gather-throughput ceiling: (gathers/s, streamed bytes/s) of a stream
    of (int32 index, FP64 value) pairs gathering from a window of x."""
    g, b = C.c_double(), C.c_double()
    call("spk_gather_rate", int(window), int(count), int(reps), C.byref(g),
         C.byref(b))
    return g.value, b.value


class DeviceVector:
    """A vector in HBM (float64 or float32)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        call("spk_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or np.asarray(a).dtype)
        v = cls(a.size, a.dtype)
        if a.size:
            call("spk_memcpy_htod", v.ptr, a.ctypes.data_as(C.c_void_p),
                 v.nbytes)
        return v

    @classmethod
    def zeros(cls, n, dtype=np.float64):
        v = cls(n, dtype)
        call("spk_memset", v.ptr, 0, max(v.nbytes, 1))
        return v

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.n:
            call("spk_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr,
                 self.nbytes)
        return out

    def copy_from(self, other: "DeviceVector"):
        assert other.n == self.n and other.dtype == self.dtype
        call("spk_memcpy_dtod", self.ptr, other.ptr, self.nbytes)

    def astype(self, dtype):
        out = DeviceVector(self.n, dtype)
        call("spk_convert", self.n, _code(self.dtype), self.ptr,
             _code(out.dtype), out.ptr)
        return out

    def __len__(self):
        return self.n

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib._lib is not None:
                _lib._lib.spk_free(self.ptr)
        except Exception:
            pass


class CsrMatrix:
    """Compressed_row_matrix<double> resident in HBM: int32 row_ptr/col_ind
    (narrowed from the reference's int64) and double values, plus the
    row-binned SpMV plan."""

    def __init__(self, handle, owned=True):
        self.h = handle
        self.owned = owned
        r, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        call("spk_csr_shape", self.h, C.byref(r), C.byref(c), C.byref(z))
        self.nrows, self.ncols, self.nnz = r.value, c.value, z.value

    @classmethod
    def from_host(cls, nrows, ncols, row_ptr, col_ind, vals):
        rp = np.ascontiguousarray(row_ptr, np.int64)
        ci = np.ascontiguousarray(col_ind, np.int64)
        v = np.ascontiguousarray(vals, np.float64)
        h = C.c_void_p()
        call("spk_csr_create", int(nrows), int(ncols), int(len(ci)),
             rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
             v.ctypes.data_as(C.c_void_p), C.byref(h))
        return cls(h)

    @classmethod
    def assemble(cls, nrows, ncols, rows, cols, vals=None):
        """Compressed_row_sparsity(Shape, indices) / make_matrix built on
        the device from (row, col[, value]) lists (last duplicate wins)."""
        r = DeviceVector.from_numpy(np.ascontiguousarray(rows, np.int64)
                                    .view(np.float64))
        c = DeviceVector.from_numpy(np.ascontiguousarray(cols, np.int64)
                                    .view(np.float64))
        v = None if vals is None else DeviceVector.from_numpy(
            np.ascontiguousarray(vals, np.float64))
        h = C.c_void_p()
        call("spk_csr_assemble", int(nrows), int(ncols), len(rows), r.ptr,
             c.ptr, None if v is None else v.ptr, C.byref(h))
        return cls(h)

    @classmethod
    def from_oracle(cls, A):
        """From any object with nrows/ncols/row_ptr/col_ind/vals."""
        return cls.from_host(A.nrows, A.ncols, A.row_ptr, A.col_ind, A.vals)

    def download(self):
        rp = np.zeros(self.nrows + 1, np.int64)
        ci = np.zeros(max(self.nnz, 1), np.int64)
        v = np.zeros(max(self.nnz, 1), np.float64)
        call("spk_csr_download", self.h, rp.ctypes.data_as(C.c_void_p),
             ci.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p))
        return rp, ci[:self.nnz], v[:self.nnz]

    def plan_info(self):
        nb, nl = C.c_int64(), C.c_int64()
        call("spk_csr_plan_info", self.h, C.byref(nb), C.byref(nl))
        return nb.value, nl.value

    def transpose(self) -> "CsrMatrix":
        """A^T on the device; its SpMV is multiply_transpose(A, x)."""
        h = C.c_void_p()
        call("spk_csr_transpose", self.h, C.byref(h))
        return CsrMatrix(h)

    @staticmethod
    def from_matrix_market(path) -> "CsrMatrix":
        """read_matrix_market(path) + to_compressed_row
        (matrix_market.hpp:43-104), assembled on the device."""
        h = C.c_void_p()
        call("spk_csr_read_matrix_market", os.fsencode(path), C.byref(h))
        return CsrMatrix(h)

    @staticmethod
    def from_matrix_market_text(text) -> "CsrMatrix":
        """The same from a string (an std::istream in the reference)."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        call("spk_csr_read_matrix_market_text", data, len(data), C.byref(h))
        return CsrMatrix(h)

    def to_matrix_market_text(self) -> str:
        """write_matrix_market(os, A) (matrix_market.hpp:123-148)."""
        n = C.c_size_t()
        call("spk_csr_write_matrix_market_text", self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        call("spk_csr_write_matrix_market_text", self.h, buf, n.value + 1,
             C.byref(n))
        return buf.raw[:n.value].decode()

    def write_matrix_market(self, path):
        """write_matrix_market(path, A) (matrix_market.hpp:150-160)."""
        call("spk_csr_write_matrix_market", self.h, os.fsencode(path))

    def symmetric_part(self) -> "CsrMatrix":
        """(M + M^T) / 2 on M's (structurally symmetric) pattern."""
        h = C.c_void_p()
        call("spk_csr_symmetric_part", self.h, C.byref(h))
        return CsrMatrix(h)

    def permute(self, perm, mode="both") -> "CsrMatrix":
        """rperm ('rows'), cperm ('cols') or dperm ('both', P A P^T) of
        permutation.hpp:41-148 (perm[old] = new), on the device."""
        m = {"rows": 1, "cols": 2, "both": 3}[mode]
        p = DeviceVector.from_numpy(np.ascontiguousarray(perm, np.int64)
                                    .view(np.float64))
        h = C.c_void_p()
        call("spk_csr_permute", self.h, p.ptr, m, C.byref(h))
        return CsrMatrix(h)

    def rcm(self) -> np.ndarray:
        """reverse_cuthill_mckee(A.sparsity()) (reordering.cpp:152-203) on
        the device: perm[old] = new (for permute(perm, 'both'))."""
        p = DeviceVector.zeros(self.nrows)
        call("spk_csr_rcm", self.h, p.ptr)
        return p.numpy().view(np.int64).copy()

    def pseudo_peripheral_node(self) -> int:
        """pseudo_peripheral_node (reordering.cpp:112-150)."""
        v = C.c_int64()
        call("spk_csr_pseudo_peripheral_node", self.h, C.byref(v))
        return v.value

    @property
    def panels(self) -> int:
        """The SpMV layout code: the number of column panels (1: plain
        store), or -(slots * 65536 + window) for window tiles."""
        p = C.c_int64()
        call("spk_csr_get_panels", self.h, C.byref(p))
        return p.value

    def set_window_tiles(self, slots: int = 16, window: int = 8192):
        """Window tiles: x windows staged in LDS (spk_csr_set_window_tiles)."""
        call("spk_csr_set_window_tiles", self.h, int(slots), int(window))

    def set_layout(self, code: int):
        """A layout by its code (panels, or a negative window-tile code)."""
        call("spk_csr_set_layout", self.h, int(code))

    def set_half_values(self, enable: bool = True):
        """FP32 SpMVs read the values rounded to FP16
        (spk_csr_set_half_values): for a preconditioner store."""
        call("spk_csr_set_half_values", self.h, 1 if enable else 0)

    def set_short_indices(self, enable: bool = True):
        """The row kernel reads 16-bit column offsets from each row block's
        first row (spk_csr_set_short_indices): banded stores only."""
        call("spk_csr_set_short_indices", self.h, 1 if enable else 0)

    def set_panels(self, panels: int):
        call("spk_csr_set_panels", self.h, int(panels))

    def tune_layout(self):
        """Times the SpMV over 1..64 column panels and keeps the fastest;
        returns (chosen, [(panels, us), ...])."""
        ch, nt = C.c_int64(), C.c_int64()
        tried = np.zeros(16, np.int64)
        us = np.zeros(16, np.float32)
        call("spk_csr_tune_layout", self.h, C.byref(ch),
             tried.ctypes.data_as(C.c_void_p), us.ctypes.data_as(C.c_void_p),
             C.byref(nt))
        return ch.value, [(int(tried[i]), float(us[i]))
                          for i in range(nt.value)]

    def spmv_dot(self, x: DeviceVector, y: DeviceVector,
                 result: DeviceVector):
        """y = A x and result[0] = <x, y> (device), one fused launch plus
        its finisher; nothing is synchronised."""
        call("spk_csr_spmv_dot", self.h, _code(x.dtype), x.ptr, y.ptr,
             result.ptr)

    def spmv(self, x: DeviceVector, y: DeviceVector):
        call("spk_csr_spmv", self.h, _code(x.dtype), x.ptr, y.ptr)

    def __del__(self):
        try:
            if (getattr(self, "h", None) and getattr(self, "owned", True)
                    and _lib._lib is not None):
                _lib._lib.spk_csr_destroy(self.h)
        except Exception:
            pass


def multiply(A: CsrMatrix, x: DeviceVector) -> DeviceVector:
    """multiply(A, x) (sparse_blas.hpp:17-34): returns a new vector."""
    y = DeviceVector(A.nrows, x.dtype)
    A.spmv(x, y)
    return y


def extract_diagonal(A: CsrMatrix) -> DeviceVector:
    d = DeviceVector(min(A.nrows, A.ncols), np.float64)
    call("spk_csr_extract_diagonal", A.h, d.ptr)
    return d


class Preconditioner:
    def __init__(self, handle, kind, n, owner=None):
        self.h = handle
        self.kind = kind
        self.n = n
        self._owner = owner  # keeps an SpMV preconditioner's store alive

    def apply(self, r: DeviceVector, z: DeviceVector):
        call("spk_precond_apply", self.h, _code(r.dtype), r.ptr, z.ptr)

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib._lib is not None:
                _lib._lib.spk_precond_destroy(self.h)
        except Exception:
            pass


class _Identity:
    """The identity slot (conjugate_gradient_test.cpp:45-47)."""
    h = None
    kind = "identity"


identity = _Identity()


def jacobi(A: CsrMatrix) -> Preconditioner:
    """jacobi(A) (jacobi.hpp:25-36); raises InvalidArgument on a zero
    diagonal entry."""
    h = C.c_void_p()
    call("spk_precond_create_jacobi", A.h, C.byref(h))
    return Preconditioner(h, "jacobi", min(A.nrows, A.ncols))


def jacobi_apply(M: Preconditioner, r: DeviceVector, z: DeviceVector):
    """jacobi_apply (jacobi.hpp:42-50)."""
    M.apply(r, z)


class Ilu0(Preconditioner):
    def factors(self):
        """(L, U) in the reference layout: ((rp, ci, v), (rp, ci, v))."""
        nl, nu = C.c_int64(), C.c_int64()
        call("spk_ilu0_factor_nnz", self.h, C.byref(nl), C.byref(nu))
        n = self.n
        lrp = np.zeros(n + 1, np.int64)
        urp = np.zeros(n + 1, np.int64)
        lci = np.zeros(nl.value, np.int64)
        uci = np.zeros(nu.value, np.int64)
        lv = np.zeros(nl.value, np.float64)
        uv = np.zeros(nu.value, np.float64)
        P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        call("spk_ilu0_download", self.h, P(lrp), P(lci), P(lv), P(urp),
             P(uci), P(uv))
        return (lrp, lci, lv), (urp, uci, uv)

    def levels(self):
        a, b = C.c_int64(), C.c_int64()
        call("spk_ilu0_levels", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    @property
    def schedule(self) -> str:
        """'levels' (one launch per dependency level) or 'sync-free'."""
        v = C.c_int()
        call("spk_ilu0_get_schedule", self.h, C.byref(v))
        return "levels" if v.value == 1 else "sync-free"

    @schedule.setter
    def schedule(self, name: str):
        call("spk_ilu0_set_schedule", self.h,
             {"auto": 0, "levels": 1, "sync-free": 2}[name])

    @property
    def jacobi_sweeps(self) -> int:
        """0: the exact triangular solves of ilu_apply; k > 0: k Jacobi
        sweeps per triangle (approximate apply, no reference counterpart)."""
        v = C.c_int()
        call("spk_ilu0_get_jacobi_sweeps", self.h, C.byref(v))
        return v.value

    @jacobi_sweeps.setter
    def jacobi_sweeps(self, k: int):
        call("spk_ilu0_set_jacobi_sweeps", self.h, int(k))

    @property
    def half_factors(self) -> bool:
        """The Jacobi-sweep apply's FP32 passes read the strict factors'
        values in FP16 (spk_ilu0_set_half_factors)."""
        v = C.c_int()
        call("spk_ilu0_get_half_factors", self.h, C.byref(v))
        return bool(v.value)

    @half_factors.setter
    def half_factors(self, enable: bool):
        call("spk_ilu0_set_half_factors", self.h, 1 if enable else 0)

    def forward(self, b: DeviceVector, x: DeviceVector):
        call("spk_ilu0_forward", self.h, _code(b.dtype), b.ptr, x.ptr)

    def backward(self, b: DeviceVector, x: DeviceVector):
        call("spk_ilu0_backward", self.h, _code(b.dtype), b.ptr, x.ptr)


def ilu0(A: CsrMatrix) -> Ilu0:
    """ilu0(A) (ilu.hpp:155-159)."""
    h = C.c_void_p()
    call("spk_precond_create_ilu0", A.h, 0, C.byref(h))
    return Ilu0(h, "ilu0", A.nrows)


def milu0(A: CsrMatrix) -> Ilu0:
    """milu0(A) (ilu.hpp:164-168)."""
    h = C.c_void_p()
    call("spk_precond_create_ilu0", A.h, 1, C.byref(h))
    return Ilu0(h, "milu0", A.nrows)


def ilu_apply(F: Ilu0, r: DeviceVector, z: DeviceVector):
    """ilu_apply (ilu.hpp:173-179)."""
    F.apply(r, z)


class InnerCg(Preconditioner):
    """FGMRES's flexible-slot preconditioner running a solve of its own:
    up to `iterations` PCG steps on A z = r with M_inner, in `precision`
    (inner_cg), or a mixed-precision solve (inner_ir)."""

    @property
    def inner_iterations(self) -> int:
        """Inner steps run by all applications so far."""
        t = C.c_int64()
        call("spk_precond_inner_cg_iterations", self.h, C.byref(t))
        return t.value


def inner_cg(A: CsrMatrix, M_inner=None, iterations: int = 200,
             tolerance: float = 0.0, precision: int = F32) -> InnerCg:
    """An inner-PCG preconditioner (spk_precond_create_inner_cg): a
    nonlinear preconditioner, for fgmres."""
    h = C.c_void_p()
    call("spk_precond_create_inner_cg", A.h, _h(M_inner), int(iterations),
         float(tolerance), int(precision), C.byref(h))
    return InnerCg(h, "inner_cg", A.nrows, owner=(A, M_inner))


def spai(M: CsrMatrix) -> Preconditioner:
    """z = M r for an explicit sparse approximate inverse M."""
    h = C.c_void_p()
    call("spk_precond_create_spmv", M.h, C.byref(h))
    return Preconditioner(h, "spai", M.nrows, owner=M)


# ---------------------------------------------------------------- learned SPAI

SPAI_NPARAM = 6345


class SpaiModel:
    """The learned sparse-approximate-inverse network (24 features -> 64 ->
    64 -> 9), trained on the device with bf16 MFMA kernels.  A new model
    predicts the Jacobi inverse (zero output layer)."""

    def __init__(self, seed: int = 0):
        self.h = C.c_void_p()
        call("spk_spai_model_create", seed, C.byref(self.h))

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(SPAI_NPARAM, np.float32)
        call("spk_spai_model_get_params", self.h, p.ctypes.data_as(C.c_void_p))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == SPAI_NPARAM
        call("spk_spai_model_set_params", self.h, p.ctypes.data_as(C.c_void_p))

    @property
    def objective(self) -> int:
        """SPAI_LOSS_INVERSE or SPAI_LOSS_SYMMETRIC (the training loss)."""
        v = C.c_int32()
        call("spk_spai_model_get_objective", self.h, C.byref(v))
        return v.value

    @objective.setter
    def objective(self, v: int):
        call("spk_spai_model_set_objective", self.h, int(v))

    def save(self, path: str):
        """The training state (weights, Adam moments, step) to a checkpoint
        file (spk_spai_model_save)."""
        call("spk_spai_model_save", self.h, os.fsencode(path))

    def load(self, path: str) -> "SpaiModel":
        """Replaces this model's training state with a checkpoint's
        (spk_spai_model_load); returns the model."""
        call("spk_spai_model_load", self.h, os.fsencode(path))
        return self

    def loss_grad(self, A: CsrMatrix):
        """((1/n)||I - M A||_F^2, gradient) for one matrix, no update."""
        loss = C.c_double()
        g = np.zeros(SPAI_NPARAM, np.float32)
        call("spk_spai_loss_grad", self.h, A.h, C.byref(loss),
             g.ctypes.data_as(C.c_void_p))
        return loss.value, g

    def train(self, mats, epochs=1, learning_rate=1e-3, batch_matrices=1,
              shuffle_seed=0) -> np.ndarray:
        """Adam over the matrices; returns each step's mean loss."""
        hs = (C.c_void_p * len(mats))(*[m.h.value for m in mats])
        cfg = _lib.SpaiTrainConfig(epochs, learning_rate, batch_matrices,
                                   shuffle_seed)
        steps = epochs * ((len(mats) + batch_matrices - 1) // batch_matrices)
        hist = np.zeros(max(steps, 1), np.float64)
        call("spk_spai_train", self.h, hs, len(mats), C.byref(cfg),
             hist.ctypes.data_as(C.c_void_p))
        return hist[:steps]

    def predict(self, A: CsrMatrix) -> CsrMatrix:
        """M on A's pattern."""
        h = C.c_void_p()
        call("spk_spai_predict", self.h, A.h, C.byref(h))
        return CsrMatrix(h)

    def preconditioner(self, A: CsrMatrix) -> Preconditioner:
        """z = M r with the predicted M (one SpMV)."""
        return spai(self.predict(A))

    def __del__(self):
        try:
            if self.h and _lib._lib is not None:
                _lib._lib.spk_spai_model_destroy(self.h)
        except Exception:
            pass


def spai_features(A: CsrMatrix) -> np.ndarray:
    """The 24 per-row features the network reads (n x 24 float)."""
    X = DeviceVector(A.nrows * 24, np.float32)
    call("spk_spai_features", A.h, X.ptr)
    return X.numpy().reshape(A.nrows, 24)


# ---------------------------------------------------------------- batched PCG

@dataclass
class BatchConfig:
    """Batched Jacobi-PCG: FP64 true-residual tolerance per system, inner
    solve precision and budget."""
    tolerance: float = 1e-8
    inner_tolerance: float = 1e-4
    max_outer: int = 30
    max_inner: int = 20000
    precision: int = F32


class BatchedSystems:
    """Many independent SPD systems (diagonal blocks of one store) solved by
    one persistent batched Jacobi-PCG launch."""

    def __init__(self, handle):
        self.h = handle
        c, t, z = C.c_int64(), C.c_int64(), C.c_int64()
        m = C.c_void_p()
        call("spk_batch_info", self.h, C.byref(c), C.byref(t), C.byref(z),
             C.byref(m))
        self.count, self.total_rows, self.nnz = c.value, t.value, z.value
        self.matrix = CsrMatrix(m.value, owned=False)

    @classmethod
    def diffusion(cls, count, seed, side_min=32, side_max=256):
        h = C.c_void_p()
        call("spk_batch_create_diffusion", count, seed, side_min, side_max,
             C.byref(h))
        return cls(h)

    @classmethod
    def from_blocks(cls, A: CsrMatrix, row_offsets):
        ro = np.ascontiguousarray(row_offsets, np.int64)
        h = C.c_void_p()
        call("spk_batch_create", A.h, len(ro) - 1,
             ro.ctypes.data_as(C.c_void_p), C.byref(h))
        return cls(h)

    def row_offsets(self):
        ro = np.zeros(self.count + 1, np.int64)
        call("spk_batch_row_offsets", self.h, ro.ctypes.data_as(C.c_void_p))
        return ro

    @property
    def tuning(self):
        t = _lib.BatchTuning()
        call("spk_batch_get_tuning", self.h, C.byref(t))
        return (t.block, t.wg_per_cu, t.lds_rows)

    @tuning.setter
    def tuning(self, v):
        t = _lib.BatchTuning(int(v[0]), int(v[1]), int(v[2]), 0)
        call("spk_batch_set_tuning", self.h, C.byref(t))

    def save_tuning(self, path: str, tried=None):
        """The batch's launch shape (as autotune chose it) to a small JSON
        file keyed by the GPU target, with the timings of the shapes tried
        (autotune's second result) when given."""
        import json
        block, wg, lds = self.tuning
        doc = {"format": "sparkit_amd batched PCG tuning", "version": 1,
               "gfx_target": _gfx_target(),
               "tuning": {"block": block, "wg_per_cu": wg, "lds_rows": lds}}
        if tried is not None:
            doc["tried"] = [{"block": t[0], "wg_per_cu": t[1], "lds_rows": t[2],
                             "ms": float(ms)} for t, ms in tried]
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)

    def load_tuning(self, path: str):
        """Applies a saved launch shape; InvalidArgument when it was tuned on
        another GPU target (or the file is not such a record)."""
        import json
        with open(path) as f:
            doc = json.load(f)
        if doc.get("format") != "sparkit_amd batched PCG tuning" or \
                doc.get("version") != 1:
            raise InvalidArgument(1, "load_tuning", f"{path}: not a tuning record")
        if doc.get("gfx_target") != _gfx_target():
            raise InvalidArgument(1, "load_tuning",
                                  f"{path}: tuned on {doc.get('gfx_target')}, "
                                  f"this GPU is {_gfx_target()}")
        t = doc["tuning"]
        self.tuning = (t["block"], t["wg_per_cu"], t["lds_rows"])
        return self.tuning

    @staticmethod
    def _cfg(cfg):
        return _lib.BatchConfig(float(cfg.tolerance), float(cfg.inner_tolerance),
                                int(cfg.max_outer), int(cfg.max_inner),
                                int(cfg.precision), 0)

    def autotune(self, cfg: BatchConfig, b, x, sample=512):
        """Times the candidate launch shapes on a sample of the systems;
        returns (chosen, [(tuning, ms), ...])."""
        c = self._cfg(cfg)
        ch = _lib.BatchTuning()
        tried = (_lib.BatchTuning * 64)()
        ms = np.zeros(64, np.float32)
        nt = C.c_int64()
        call("spk_batch_autotune", self.h, C.byref(c), b.ptr, x.ptr, sample,
             C.byref(ch), tried, ms.ctypes.data_as(C.c_void_p), C.byref(nt))
        lst = [((tried[i].block, tried[i].wg_per_cu, tried[i].lds_rows),
                float(ms[i])) for i in range(nt.value)]
        return (ch.block, ch.wg_per_cu, ch.lds_rows), lst

    def solve(self, cfg: BatchConfig, b, x, use_graph=False, summary=True):
        c = self._cfg(cfg)
        s = _lib.BatchSummary()
        call("spk_batch_pcg_solve", self.h, C.byref(c), b.ptr, x.ptr,
             int(use_graph), C.byref(s) if summary else None)
        if not summary:
            return None
        return {"total_inner_iterations": s.total_inner_iterations,
                "max_inner_iterations": s.max_inner_iterations,
                "total_outer_iterations": s.total_outer_iterations,
                "converged": s.converged, "max_relres": s.max_relres}

    def results(self):
        it = np.zeros(self.count, np.int32)
        ou = np.zeros(self.count, np.int32)
        rr = np.zeros(self.count, np.float64)
        cv = np.zeros(self.count, np.int32)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        call("spk_batch_results", self.h, P(it), P(ou), P(rr), P(cv))
        return it, ou, rr, cv

    def __del__(self):
        try:
            if self.h and _lib._lib is not None:
                _lib._lib.spk_batch_destroy(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------- solvers

@dataclass
class CGConfig:
    """CGConfig (conjugate_gradient.hpp:47-74) + the arithmetic precision of
    the Krylov vectors."""
    tolerance: float = 0.0
    restart_iterations: int = 0
    max_iterations: int = 0
    collect_residuals: bool = False
    precision: int = F64


@dataclass
class GmresConfig:
    """Gmres_config (gmres.hpp:24-47) + the arithmetic precision of the
    Krylov vectors and the orthogonalization (ORTHO_MGS, the reference's,
    ORTHO_CGS2, or ORTHO_CGS1 where a true residual decides)."""
    tolerance: float = 0.0
    restart_dimension: int = 0
    max_iterations: int = 0
    collect_residuals: bool = False
    precision: int = F64
    orthogonalization: int = 0


FgmresConfig = GmresConfig  # Fgmres_config has the same fields


@dataclass
class BicgstabConfig:
    """Bicgstab_config (bicgstab.hpp:24-42)."""
    tolerance: float = 0.0
    max_iterations: int = 0
    collect_residuals: bool = False
    precision: int = F64


@dataclass
class Summary:
    """CGSummary / Gmres_summary / Fgmres_summary / Bicgstab_summary."""
    residual_norm: float = 0.0
    computed_iterations: int = 0
    converged: bool = False
    iteration_residuals: list = field(default_factory=list)
    breakdown: bool = False


def _cfg(c, restart):
    return _lib.SolverConfig(float(c.tolerance), int(restart),
                             int(c.max_iterations),
                             1 if c.collect_residuals else 0,
                             int(c.precision),
                             int(getattr(c, "orthogonalization", 0)), 0)


def _run(name, args, cfg):
    s = _lib.Summary()
    hist = np.zeros(max(cfg.max_iterations + 1, 1), np.float64)
    call(name, *args, C.byref(s), hist.ctypes.data_as(C.c_void_p))
    return Summary(s.residual_norm, s.computed_iterations, bool(s.converged),
                   hist[:s.history_len].tolist(), s.status == BREAKDOWN)


def _h(M):
    return None if M is None else M.h


def _vec_check(A, b, x, precision):
    dt = np.float32 if precision == F32 else np.float64
    if b.dtype != dt or x.dtype != dt:
        raise InvalidArgument(1, "solver", "b and x must match the precision")
    if b.n != A.nrows or x.n != A.nrows:
        raise InvalidArgument(1, "solver", "vector size mismatch")


def conjugate_gradient(b, x, cfg: CGConfig, A: CsrMatrix, M_left=identity,
                       M_right=identity) -> Summary:
    """conjugate_gradient() (conjugate_gradient.hpp:339-366); x updated in
    place."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, cfg.restart_iterations)
    return _run("spk_cg_solve", (A.h, _h(M_left), _h(M_right), C.byref(c),
                                 b.ptr, x.ptr), cfg)


def conjugate_gradient_noalloc(b, x, cfg: CGConfig, A: CsrMatrix,
                               M_left=identity, M_right=identity) -> Summary:
    """conjugate_gradient() through spk_cg_solve_noalloc: the same solve,
    with every work buffer in the library's static device scratch (no
    device allocation; for callers that cannot allocate)."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, cfg.restart_iterations)
    return _run("spk_cg_solve_noalloc", (A.h, _h(M_left), _h(M_right),
                                         C.byref(c), b.ptr, x.ptr), cfg)


def gmres(b, x, cfg: GmresConfig, A: CsrMatrix, M_left=identity,
          M_right=identity) -> Summary:
    """gmres() (gmres.hpp:400-427)."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, cfg.restart_dimension)
    return _run("spk_gmres_solve", (A.h, _h(M_left), _h(M_right), C.byref(c),
                                    b.ptr, x.ptr), cfg)


def fgmres(b, x, cfg: GmresConfig, A: CsrMatrix, M=identity) -> Summary:
    """fgmres() (fgmres.hpp:452-470)."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, cfg.restart_dimension)
    return _run("spk_fgmres_solve", (A.h, _h(M), C.byref(c), b.ptr, x.ptr),
                cfg)


CgsConfig = BicgstabConfig  # Cgs_config has the same fields (cgs.hpp:24-42)
BicgConfig = BicgstabConfig  # Bicg_config has the same fields (bicg.hpp:24-42)


def bicg(b, x, cfg: BicgstabConfig, A: CsrMatrix, At: CsrMatrix,
         M_left=identity, M_right=identity) -> Summary:
    """bicg() (bicg.hpp:86-397) with the transpose operator given as the
    store At = A.transpose(); counts matrix-vector products."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, 0)
    return _run("spk_bicg_solve", (A.h, At.h, _h(M_left), _h(M_right),
                                   C.byref(c), b.ptr, x.ptr), cfg)


def cgs(b, x, cfg: BicgstabConfig, A: CsrMatrix, M_left=identity,
        M_right=identity) -> Summary:
    """cgs() (cgs.hpp:383-395); computed_iterations counts matrix-vector
    products, two per step."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, 0)
    return _run("spk_cgs_solve", (A.h, _h(M_left), _h(M_right), C.byref(c),
                                  b.ptr, x.ptr), cfg)


MinresConfig = BicgstabConfig  # Minres_config (minres.hpp:24-45)
QmrConfig = BicgstabConfig  # Qmr_config (qmr.hpp:24-42)


def minres(b, x, cfg: BicgstabConfig, A: CsrMatrix, M_left=identity,
           M_right=identity) -> Summary:
    """minres() (minres.hpp:364-395) for symmetric A and M = M_L M_R."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, 0)
    return _run("spk_minres_solve", (A.h, _h(M_left), _h(M_right), C.byref(c),
                                     b.ptr, x.ptr), cfg)


def qmr(b, x, cfg: BicgstabConfig, A: CsrMatrix, At: CsrMatrix,
        M_left=identity, M_right=identity) -> Summary:
    """qmr() (qmr.hpp:455-489) with the transpose operator given as the
    store At = A.transpose(); counts matrix-vector products."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, 0)
    return _run("spk_qmr_solve", (A.h, At.h, _h(M_left), _h(M_right),
                                  C.byref(c), b.ptr, x.ptr), cfg)


def bicgstab(b, x, cfg: BicgstabConfig, A: CsrMatrix, M_left=identity,
             M_right=identity) -> Summary:
    """bicgstab() (bicgstab.hpp:383-410); computed_iterations counts
    matrix-vector products."""
    _vec_check(A, b, x, cfg.precision)
    c = _cfg(cfg, 0)
    return _run("spk_bicgstab_solve", (A.h, _h(M_left), _h(M_right),
                                       C.byref(c), b.ptr, x.ptr), cfg)


IR_GMRES, IR_FGMRES, IR_CG, IR_BICGSTAB, IR_CG_RELIABLE = 0, 1, 2, 3, 4


@dataclass
class IrConfig:
    """Mixed-precision iterative refinement: FP64 true residual outside,
    FP32 Krylov inside (FP64 dots/scalars)."""
    tolerance: float = 1e-10
    max_outer: int = 20
    inner_tolerance: float = 1e-4
    restart: int = 50
    max_inner: int = 500
    inner_method: int = IR_GMRES
    orthogonalization: int = 0


@dataclass
class IrSummary:
    relres: float
    outer_iterations: int
    inner_iterations: int
    converged: bool
    relres_history: list


def ir_solve(b: DeviceVector, x: DeviceVector, cfg: IrConfig, A: CsrMatrix,
             M=identity) -> IrSummary:
    """x (FP64, updated in place) from b (FP64) by iterative refinement."""
    if b.dtype != np.float64 or x.dtype != np.float64:
        raise InvalidArgument(1, "ir_solve", "b and x must be float64")
    c = _ir_struct(cfg)
    s = _lib.IrSummary()
    hist = np.zeros(cfg.max_outer + 2, np.float64)
    call("spk_ir_solve", A.h, _h(M), C.byref(c), b.ptr, x.ptr, C.byref(s),
         hist.ctypes.data_as(C.c_void_p))
    return IrSummary(s.relres, s.outer_iterations, s.inner_iterations,
                     bool(s.converged), hist[:s.outer_iterations + 1].tolist())


def _ir_struct(cfg: IrConfig):
    return _lib.IrConfig(float(cfg.tolerance), int(cfg.max_outer),
                         float(cfg.inner_tolerance), int(cfg.restart),
                         int(cfg.max_inner), int(cfg.inner_method),
                         int(cfg.orthogonalization))


def inner_ir(A: CsrMatrix, M_inner=None, cfg: IrConfig = None) -> InnerCg:
    """FGMRES's flexible-slot preconditioner running a mixed-precision
    solve (spk_precond_create_inner_ir): z = ir_solve(A z = r) from z = 0
    with M_inner; inner_method IR_CG, IR_CG_RELIABLE or IR_BICGSTAB.
    inner_iterations counts the FP32 steps of all applications."""
    if cfg is None:
        cfg = IrConfig(tolerance=1e-4, max_outer=50, inner_tolerance=1e-1,
                       restart=0, max_inner=100000,
                       inner_method=IR_CG_RELIABLE)
    c = _ir_struct(cfg)
    h = C.c_void_p()
    call("spk_precond_create_inner_ir", A.h, _h(M_inner), C.byref(c),
         C.byref(h))
    return InnerCg(h, "inner_ir", A.nrows, owner=(A, M_inner))


FALLBACK_AUTO, FALLBACK_JACOBI_PCG, FALLBACK_ILU_GMRES, FALLBACK_NONE = 0, 1, 2, 3
METHODS = {0: "learned_fgmres", 1: "jacobi_pcg", 2: "ilu_gmres"}
REASONS = {1: "no_preconditioner", 2: "bad_values", 4: "bad_diagonal",
           8: "stalled", 16: "not_converged"}


@dataclass
class LearnedConfig:
    """spk_learned_config: FGMRES(restart) with the learned M, guarded; the
    fallback (Jacobi-PCG for symmetric A, else ILU(0)-GMRES) when M is bad,
    FGMRES stalls or the FP64 true residual misses the tolerance."""
    tolerance: float = 1e-8
    restart: int = 30
    max_iterations: int = 50000
    fallback_max_iterations: int = 50000
    stall_factor: float = 0.999
    stall_cycles: int = 20
    fallback: int = FALLBACK_AUTO
    orthogonalization: int = ORTHO_CGS2


@dataclass
class LearnedSummary:
    method: str
    reasons: list
    converged: bool
    learned_iterations: int
    fallback_iterations: int
    bad_values: int
    bad_diagonal: int
    learned_relres: float
    relres: float

    @property
    def fell_back(self) -> bool:
        return self.method != "learned_fgmres"

    @property
    def computed_iterations(self) -> int:
        """FGMRES and fallback iterations together."""
        return self.learned_iterations + self.fallback_iterations


def learned_solve(b: DeviceVector, x: DeviceVector, cfg: LearnedConfig,
                  A: CsrMatrix, M=None) -> LearnedSummary:
    """x (FP64, in place) by FGMRES with the explicit approximate inverse M
    (an SpMV preconditioner, e.g. SpaiModel.preconditioner(A)) and the
    fallback; the FP64 true residual decides."""
    if b.dtype != np.float64 or x.dtype != np.float64:
        raise InvalidArgument(1, "learned_solve", "b and x must be float64")
    c = _lib.LearnedConfig(float(cfg.tolerance), int(cfg.restart),
                           int(cfg.max_iterations),
                           int(cfg.fallback_max_iterations),
                           float(cfg.stall_factor), int(cfg.stall_cycles),
                           int(cfg.fallback), int(cfg.orthogonalization), 0)
    s = _lib.LearnedSummary()
    call("spk_learned_solve", A.h, _h(M), C.byref(c), b.ptr, x.ptr,
         C.byref(s))
    return LearnedSummary(
        METHODS[s.method], [v for k, v in REASONS.items() if s.reason & k],
        bool(s.converged), s.learned_iterations, s.fallback_iterations,
        s.bad_values, s.bad_diagonal, s.learned_relres, s.relres)


def dot(x: DeviceVector, y: DeviceVector) -> float:
    """<x, y> summed in FP64 on the device, in a fixed order (the
    std::inner_product of conjugate_gradient.hpp:194,236; the order of the
    sum differs, so results agree to rounding, not bit for bit)."""
    if x.n != y.n or x.dtype != y.dtype:
        raise InvalidArgument(1, "dot", "x and y differ in length or dtype")
    r = C.c_double()
    call("spk_dot", _code(x.dtype), x.n, x.ptr, y.ptr, C.byref(r))
    return r.value


def axpby(a: float, x: DeviceVector, b: float, y: DeviceVector):
    """y = a x + b y in the vectors' precision, in place."""
    if x.n != y.n or x.dtype != y.dtype:
        raise InvalidArgument(1, "axpby", "x and y differ in length or dtype")
    call("spk_axpby", _code(x.dtype), x.n, float(a), x.ptr, float(b), y.ptr)


def true_relres(A: CsrMatrix, b: DeviceVector, x: DeviceVector) -> float:
    """||b - A x|| / ||b|| in FP64 on the device."""
    r = C.c_double()
    call("spk_true_relres", A.h, b.ptr, x.ptr, C.byref(r))
    return r.value


# ------------------------------------------------------------- generators

def _gen(name, *args):
    h = C.c_void_p()
    call(name, *args, C.byref(h))
    return CsrMatrix(h)


def poisson_2d(nx, ny=None, shift=0.0) -> CsrMatrix:
    """poisson_2d (matgen.hpp:173-265) built on the device."""
    return _gen("spk_gen_poisson_2d", nx, nx if ny is None else ny,
                float(shift))


def poisson_3d(nx, ny=None, nz=None) -> CsrMatrix:
    """7-point Dirichlet Laplacian (diagonal 6) built on the device."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _gen("spk_gen_poisson_3d", nx, ny, nz)


def convdiff_3d(nx, ny, nz, d, cx, cy, cz) -> CsrMatrix:
    """7-point upwind convection-diffusion, Dirichlet diagonal."""
    return _gen("spk_gen_convdiff_3d", nx, ny, nz, float(d), float(cx),
                float(cy), float(cz))


def powerlaw_spd(n: int, seed: int) -> CsrMatrix:
    """Power-law SPD matrix (Zipf(1.5) row lengths, symmetrised,
    diagonally dominant) built on the device."""
    return _gen("spk_gen_powerlaw_spd", n, seed)


def aniso_q1_2d(n_side: int, seed: int, sigma: float = 1.5,
                isotropic: bool = False) -> CsrMatrix:
    """2D anisotropic variable-coefficient diffusion, bilinear finite
    elements on n_side x n_side interior nodes (9-point SPD), conductivity
    lognormal(0, sigma), anisotropy log-uniform [1, 1e3], random rotation."""
    return _gen("spk_gen_aniso_q1_2d", n_side, seed, float(sigma),
                int(isotropic))


def uniform_pm1(seed: int, n: int) -> DeviceVector:
    """splitmix64 uniform[-1,1] vector generated on the device."""
    v = DeviceVector(n, np.float64)
    call("spk_fill_uniform_pm1", seed, n, v.ptr)
    return v
