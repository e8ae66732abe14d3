"""Loader and ctypes declarations of libsparkit_hip.so (the C ABI declared in
include/sparkit_hip.h).  There is no fallback: if the library is missing or
cannot be loaded, every entry point raises."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparkit_hip.so")

# status codes (include/sparkit_hip.h)
OK = 0
INVALID_ARGUMENT = 1
LOGIC_ERROR = 2
OUT_OF_MEMORY = 3
HIP_ERROR = 4
BREAKDOWN = 5
NO_DEVICE = 6
RUNTIME_ERROR = 7

F64 = 0
F32 = 1

ORTHO_MGS = 0
ORTHO_CGS2 = 1
ORTHO_CGS1 = 2
SPAI_LOSS_INVERSE = 0    # (1/n) ||I - M A||_F^2
SPAI_LOSS_SYMMETRIC = 1  # (1/n) ||I - S A||_F^2, S = (M + M^T) / 2


class SpkError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


class InvalidArgument(SpkError, ValueError):
    """std::invalid_argument in the reference."""


class LogicError(SpkError):
    """std::logic_error in the reference."""


class ReadError(SpkError):
    """std::runtime_error in the reference (matrix_market.hpp)."""


class SolverConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("restart", C.c_int64),
                ("max_iterations", C.c_int64),
                ("collect_residuals", C.c_int32), ("precision", C.c_int32),
                ("orthogonalization", C.c_int32), ("reserved", C.c_int32)]


class Summary(C.Structure):
    _fields_ = [("residual_norm", C.c_double),
                ("computed_iterations", C.c_int64),
                ("converged", C.c_int32), ("status", C.c_int32),
                ("history_len", C.c_int64)]


class IrConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_outer", C.c_int64),
                ("inner_tolerance", C.c_double), ("restart", C.c_int64),
                ("max_inner", C.c_int64), ("inner_method", C.c_int32),
                ("orthogonalization", C.c_int32)]


class SpaiTrainConfig(C.Structure):
    _fields_ = [("epochs", C.c_int64), ("learning_rate", C.c_double),
                ("batch_matrices", C.c_int64), ("shuffle_seed", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("inner_tolerance", C.c_double),
                ("max_outer", C.c_int64), ("max_inner", C.c_int64),
                ("precision", C.c_int32), ("pad", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("wg_per_cu", C.c_int32),
                ("lds_rows", C.c_int32), ("pad", C.c_int32)]


class BatchSummary(C.Structure):
    _fields_ = [("total_inner_iterations", C.c_int64),
                ("max_inner_iterations", C.c_int64),
                ("total_outer_iterations", C.c_int64),
                ("converged", C.c_int64), ("max_relres", C.c_double)]


class IrSummary(C.Structure):
    _fields_ = [("relres", C.c_double), ("outer_iterations", C.c_int64),
                ("inner_iterations", C.c_int64), ("converged", C.c_int32),
                ("status", C.c_int32)]


class LearnedConfig(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("restart", C.c_int64),
                ("max_iterations", C.c_int64),
                ("fallback_max_iterations", C.c_int64),
                ("stall_factor", C.c_double), ("stall_cycles", C.c_int32),
                ("fallback", C.c_int32), ("orthogonalization", C.c_int32),
                ("pad", C.c_int32)]


class LearnedSummary(C.Structure):
    _fields_ = [("method", C.c_int32), ("reason", C.c_int32),
                ("converged", C.c_int32), ("pad", C.c_int32),
                ("learned_iterations", C.c_int64),
                ("fallback_iterations", C.c_int64),
                ("bad_values", C.c_int64), ("bad_diagonal", C.c_int64),
                ("learned_relres", C.c_double), ("relres", C.c_double)]


vp = C.c_void_p
i64 = C.c_int64
u64 = C.c_uint64
dbl = C.c_double
i32 = C.c_int
P_i64 = C.POINTER(C.c_int64)
P_dbl = C.POINTER(C.c_double)

# name: (argtypes, restype)  -- restype None means spk_status (int)
_SIGS = {
    "spk_last_error": ([], C.c_char_p),
    "spk_init": ([i32], None),
    "spk_finalize": ([], None),
    "spk_set_stream": ([vp], None),
    "spk_get_stream": ([], vp),
    "spk_synchronize": ([], None),
    "spk_device_info": ([C.POINTER(C.c_int), C.c_char_p, C.c_size_t], None),
    "spk_malloc": ([C.POINTER(vp), C.c_size_t], None),
    "spk_free": ([vp], None),
    "spk_memcpy_htod": ([vp, vp, C.c_size_t], None),
    "spk_memcpy_dtoh": ([vp, vp, C.c_size_t], None),
    "spk_memcpy_dtod": ([vp, vp, C.c_size_t], None),
    "spk_memset": ([vp, i32, C.c_size_t], None),
    "spk_timer_start": ([], None),
    "spk_timer_stop": ([C.POINTER(C.c_float)], None),
    "spk_gather_rate": ([C.c_int64, C.c_int64, C.c_int,
                         C.POINTER(C.c_double), C.POINTER(C.c_double)], None),
    "spk_csr_create": ([i64, i64, i64, vp, vp, vp, C.POINTER(vp)], None),
    "spk_csr_create_device": ([i64, i64, i64, vp, vp, vp, C.POINTER(vp)],
                              None),
    "spk_csr_destroy": ([vp], None),
    "spk_csr_shape": ([vp, P_i64, P_i64, P_i64], None),
    "spk_csr_download": ([vp, vp, vp, vp], None),
    "spk_csr_device_arrays": ([vp, i32, C.POINTER(vp), C.POINTER(vp),
                               C.POINTER(vp)], None),
    "spk_csr_plan_info": ([vp, P_i64, P_i64], None),
    "spk_csr_spmv": ([vp, i32, vp, vp], None),
    "spk_csr_spmv_dot": ([vp, i32, vp, vp, vp], None),
    "spk_csr_extract_diagonal": ([vp, vp], None),
    "spk_csr_assemble": ([i64, i64, i64, vp, vp, vp, C.POINTER(vp)], None),
    "spk_csr_permute": ([vp, vp, i32, C.POINTER(vp)], None),
    "spk_csr_set_panels": ([vp, i64], None),
    "spk_csr_set_window_tiles": ([vp, i64, i64], None),
    "spk_csr_set_half_values": ([vp, i32], None),
    "spk_csr_set_short_indices": ([vp, i32], None),
    "spk_ilu0_set_half_factors": ([vp, i32], None),
    "spk_ilu0_get_half_factors": ([vp, C.POINTER(C.c_int)], None),
    "spk_csr_set_layout": ([vp, i64], None),
    "spk_csr_get_panels": ([vp, P_i64], None),
    "spk_csr_tune_layout": ([vp, P_i64, vp, vp, P_i64], None),
    "spk_dot": ([i32, i64, vp, vp, P_dbl], None),
    "spk_precond_create_inner_cg": ([vp, vp, i64, dbl, i32, C.POINTER(vp)],
                                    None),
    "spk_precond_inner_cg_iterations": ([vp, P_i64], None),
    "spk_precond_create_inner_ir": ([vp, vp, C.POINTER(IrConfig), C.POINTER(vp)],
                                    None),
    "spk_learned_solve": ([vp, vp, vp, vp, vp, vp], None),
    "spk_axpby": ([i32, i64, dbl, vp, dbl, vp], None),
    "spk_convert": ([i64, i32, vp, i32, vp], None),
    "spk_precond_create_jacobi": ([vp, C.POINTER(vp)], None),
    "spk_precond_create_ilu0": ([vp, i32, C.POINTER(vp)], None),
    "spk_precond_create_identity": ([i64, C.POINTER(vp)], None),
    "spk_precond_create_spmv": ([vp, C.POINTER(vp)], None),
    "spk_precond_destroy": ([vp], None),
    "spk_precond_apply": ([vp, i32, vp, vp], None),
    "spk_ilu0_factor_nnz": ([vp, P_i64, P_i64], None),
    "spk_ilu0_download": ([vp, vp, vp, vp, vp, vp, vp], None),
    "spk_ilu0_levels": ([vp, P_i64, P_i64], None),
    "spk_ilu0_set_schedule": ([vp, i32], None),
    "spk_ilu0_get_schedule": ([vp, C.POINTER(C.c_int)], None),
    "spk_ilu0_set_jacobi_sweeps": ([vp, i32], None),
    "spk_ilu0_get_jacobi_sweeps": ([vp, C.POINTER(C.c_int)], None),
    "spk_ilu0_forward": ([vp, i32, vp, vp], None),
    "spk_ilu0_backward": ([vp, i32, vp, vp], None),
    "spk_cg_solve": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                      C.POINTER(Summary), vp], None),
    "spk_cg_solve_noalloc": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                              C.POINTER(Summary), vp], None),
    "spk_gmres_solve": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                         C.POINTER(Summary), vp], None),
    "spk_fgmres_solve": ([vp, vp, C.POINTER(SolverConfig), vp, vp,
                          C.POINTER(Summary), vp], None),
    "spk_bicgstab_solve": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                            C.POINTER(Summary), vp], None),
    "spk_cgs_solve": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                       C.POINTER(Summary), vp], None),
    "spk_bicg_solve": ([vp, vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                        C.POINTER(Summary), vp], None),
    "spk_minres_solve": ([vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                          C.POINTER(Summary), vp], None),
    "spk_qmr_solve": ([vp, vp, vp, vp, C.POINTER(SolverConfig), vp, vp,
                       C.POINTER(Summary), vp], None),
    "spk_csr_transpose": ([vp, C.POINTER(vp)], None),
    "spk_csr_rcm": ([vp, vp], None),
    "spk_csr_pseudo_peripheral_node": ([vp, C.POINTER(C.c_int64)], None),
    "spk_csr_read_matrix_market": ([C.c_char_p, C.POINTER(vp)], None),
    "spk_csr_read_matrix_market_text": ([C.c_char_p, C.c_size_t, C.POINTER(vp)],
                                        None),
    "spk_csr_write_matrix_market": ([vp, C.c_char_p], None),
    "spk_csr_write_matrix_market_text": ([vp, C.c_char_p, C.c_size_t,
                                          C.POINTER(C.c_size_t)], None),
    "spk_csr_symmetric_part": ([vp, C.POINTER(vp)], None),
    "spk_true_relres": ([vp, vp, vp, P_dbl], None),
    "spk_ir_solve": ([vp, vp, C.POINTER(IrConfig), vp, vp,
                      C.POINTER(IrSummary), vp], None),
    "spk_fill_uniform_pm1": ([u64, i64, vp], None),
    "spk_gen_poisson_2d": ([i64, i64, dbl, C.POINTER(vp)], None),
    "spk_gen_poisson_3d": ([i64, i64, i64, C.POINTER(vp)], None),
    "spk_gen_convdiff_3d": ([i64, i64, i64, dbl, dbl, dbl, dbl,
                             C.POINTER(vp)], None),
    "spk_gen_powerlaw_spd": ([i64, u64, C.POINTER(vp)], None),
    "spk_gen_aniso_q1_2d": ([i64, u64, dbl, i32, C.POINTER(vp)], None),
    "spk_batch_create_diffusion": ([i64, u64, i64, i64, C.POINTER(vp)], None),
    "spk_batch_create": ([vp, i64, vp, C.POINTER(vp)], None),
    "spk_batch_destroy": ([vp], None),
    "spk_batch_info": ([vp, P_i64, P_i64, P_i64, C.POINTER(vp)], None),
    "spk_batch_row_offsets": ([vp, vp], None),
    "spk_batch_set_tuning": ([vp, C.POINTER(BatchTuning)], None),
    "spk_batch_get_tuning": ([vp, C.POINTER(BatchTuning)], None),
    "spk_batch_autotune": ([vp, C.POINTER(BatchConfig), vp, vp, i64,
                            C.POINTER(BatchTuning), vp, vp, P_i64], None),
    "spk_batch_pcg_solve": ([vp, C.POINTER(BatchConfig), vp, vp, i32,
                             C.POINTER(BatchSummary)], None),
    "spk_batch_results": ([vp, vp, vp, vp, vp], None),
    "spk_spai_model_create": ([u64, C.POINTER(vp)], None),
    "spk_spai_model_destroy": ([vp], None),
    "spk_spai_model_get_params": ([vp, vp], None),
    "spk_spai_model_set_params": ([vp, vp], None),
    "spk_spai_model_set_objective": ([vp, i32], None),
    "spk_spai_model_get_objective": ([vp, C.POINTER(i32)], None),
    "spk_spai_model_save": ([vp, C.c_char_p], None),
    "spk_spai_model_load": ([vp, C.c_char_p], None),
    "spk_spai_features": ([vp, vp], None),
    "spk_spai_loss_grad": ([vp, vp, P_dbl, vp], None),
    "spk_spai_train": ([vp, vp, i64, C.POINTER(SpaiTrainConfig), vp], None),
    "spk_spai_predict": ([vp, vp, C.POINTER(vp)], None),
}

_lib = None


def exported_names():
    return sorted(_SIGS)


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built: run __graft_entry__.build() (or make -C "
            "sparkit_amd/csrc); the HIP path has no fallback")
    L = C.CDLL(path)
    for name, (args, res) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int if res is None else res
    _lib = L
    return L


def check(status: int, where: str):
    if status == OK:
        return
    msg = load().spk_last_error().decode(errors="replace")
    if status == INVALID_ARGUMENT:
        raise InvalidArgument(status, where, msg)
    if status == LOGIC_ERROR:
        raise LogicError(status, where, msg)
    if status == RUNTIME_ERROR:
        raise ReadError(status, where, msg)
    raise SpkError(status, where, msg)


def call(name: str, *args):
    st = getattr(load(), name)(*args)
    check(st, name)
