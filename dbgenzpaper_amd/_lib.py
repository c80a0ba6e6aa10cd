"""ctypes binding of libdbgenz.so (the C ABI declared in include/dbgenz.h).

The library is built in-tree by ``dbgenzpaper_amd/build.py``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdbgenz.so")

DBG_OK = 0
DBG_ERR_ARG = 1
DBG_ERR_RANK = 2
DBG_ERR_DEVICE = 3
DBG_ERR_NONFINITE = 4
DBG_ERR_STATE = 5

PREC_FP64, PREC_MIXED = 0, 1
GEN_RICHTMYER, GEN_SOBOL = 0, 1
PHI_ERF, PHI_ERFC, PHI_TAIL = 0, 1, 2
KIND_DENSE, KIND_SPARSE, KIND_SPARSEQMC = 0, 1, 2
CONTRACT_AUTO, CONTRACT_FP32, CONTRACT_SPLIT16 = 0, 1, 2
COV_FIXED_DENSE, COV_FIXED_SPARSE, COV_MATERN, COV_LOWRANK = 0, 1, 2, 3

# every exported symbol of include/dbgenz.h
EXPORTS = (
    "dbg_last_error", "dbg_abi_version", "dbg_opts_default", "dbg_chol_genz_f64", "dbg_chol_rowmax_f64",
    "dbg_adapt_low_rank_f64", "dbg_cov_build",
    "dbg_chol_classic_f64", "dbg_plan_create", "dbg_plan_destroy", "dbg_plan_set_shifts",
    "dbg_plan_set_sobol_directions", "dbg_qmc_pnorm", "dbg_qmc_loop", "dbg_plan_sync",
    "dbg_plan_last_kernel_ms", "dbg_plan_summary", "dbg_plan_get_factor", "dbg_plan_get_reps",
    "dbg_plan_get_state", "dbg_plan_set_state", "dbg_plan_refactor", "dbg_batch_create",
    "dbg_batch_run", "dbg_batch_last_ms", "dbg_batch_tuning", "dbg_batch_set_width", "dbg_batch_destroy",
    "dbg_qmc_grad",
    "dbg_plan_refine_info", "dbg_chol_genz_f64_dev_noalloc", "dbg_plan_set_grad_budget", "dbg_plan_grad_info",
    "dbg_plan_lowrank_info",
)


class DbgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dbgenz error {code}] {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [
        ("chol_block_size", C.c_int32), ("chol_block_size2", C.c_int32),
        ("chol_block_size3", C.c_int32), ("chol_block_size4", C.c_int32),
        ("m", C.c_int64), ("block_size_i", C.c_int32), ("block_size_i2", C.c_int32),
        ("block_size_j", C.c_int32), ("n_blocks", C.c_int32), ("n_reps", C.c_int32),
        ("n_bits", C.c_int32), ("max_pts", C.c_int64), ("max_abs_err", C.c_double),
        ("precision", C.c_int32), ("generator", C.c_int32), ("phi_form", C.c_int32),
        ("kind", C.c_int32), ("sobol_skip", C.c_int64), ("contraction", C.c_int32),
        ("refine_tiles", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [("value", C.c_double), ("log_value", C.c_double), ("err3", C.c_double),
                ("rel_err3", C.c_double), ("n_pts", C.c_int64), ("rounds", C.c_int32)]


_lib = None


def lib():
    """Load libdbgenz.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with dbgenzpaper_amd/build.py "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    dp, ip32, i64p, u32p = (C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                            C.POINTER(C.c_uint32))
    vp = C.c_void_p
    L.dbg_last_error.restype = C.c_char_p
    L.dbg_abi_version.restype = C.c_int32
    L.dbg_opts_default.argtypes = [C.POINTER(Opts), C.c_int32]
    L.dbg_opts_default.restype = None
    chol = [dp, C.c_int64, C.c_int32, dp, dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.dbg_chol_genz_f64.argtypes = chol + [ip32, dp, ip32]
    L.dbg_chol_rowmax_f64.argtypes = chol + [ip32, dp, ip32]
    L.dbg_adapt_low_rank_f64.argtypes = [dp, C.c_int64, C.c_int32, dp, dp, C.c_int32, ip32, C.c_double, ip32, ip32]
    L.dbg_cov_build.argtypes = [C.c_int32, C.c_int32, dp, C.c_int32, vp, vp, C.c_int32, vp, C.c_int64, C.c_int32]
    L.dbg_chol_classic_f64.argtypes = chol + [dp]
    L.dbg_chol_genz_f64_dev_noalloc.argtypes = [vp, C.c_int64, C.c_int32, vp, vp, C.c_int32, C.c_int32, vp, vp, vp,
                                                vp]
    L.dbg_plan_create.argtypes = [C.POINTER(vp), C.POINTER(Opts), C.c_int32, dp, C.c_int64, dp, dp]
    L.dbg_plan_destroy.argtypes = [vp]
    L.dbg_plan_set_shifts.argtypes = [vp, vp]
    L.dbg_plan_set_sobol_directions.argtypes = [vp, u32p]
    L.dbg_qmc_pnorm.argtypes = [vp, C.POINTER(Result)]
    L.dbg_qmc_loop.argtypes = [vp, C.c_int64]
    L.dbg_plan_sync.argtypes = [vp]
    L.dbg_plan_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_float)]
    L.dbg_plan_summary.argtypes = [vp, C.POINTER(Result)]
    L.dbg_plan_get_factor.argtypes = [vp, dp, ip32, dp, dp, dp, ip32]
    L.dbg_plan_get_reps.argtypes = [vp, dp, ip32]
    L.dbg_plan_get_state.argtypes = [vp, i64p]
    L.dbg_plan_set_state.argtypes = [vp, C.c_int64]
    L.dbg_plan_refactor.argtypes = [vp]
    L.dbg_plan_refine_info.argtypes = [vp, ip32, dp, C.POINTER(C.c_float)]
    L.dbg_plan_set_grad_budget.argtypes = [vp, C.c_int64]
    L.dbg_plan_grad_info.argtypes = [vp, i64p, ip32, i64p]
    L.dbg_plan_lowrank_info.argtypes = [vp, ip32, ip32, ip32]
    L.dbg_batch_create.argtypes = [C.POINTER(vp), C.POINTER(Opts), C.c_int32, ip32, dp, dp, dp, dp]
    L.dbg_batch_run.argtypes = [vp, dp]
    L.dbg_batch_last_ms.argtypes = [vp, C.POINTER(C.c_float)]
    L.dbg_batch_tuning.argtypes = [vp, ip32, C.POINTER(C.c_float)]
    L.dbg_batch_set_width.argtypes = [vp, C.c_int32]
    L.dbg_batch_destroy.argtypes = [vp]
    for name in EXPORTS:
        if name not in ("dbg_last_error", "dbg_abi_version", "dbg_opts_default"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != DBG_OK:
        msg = lib().dbg_last_error().decode(errors="replace")
        raise DbgError(rc, msg)


def ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))
