"""ctypes binding of libmfla_hip.so (include/mfla_hip.h).

The library is the product: every numerical operation of this package runs
in it on the GPU. If it is missing or no GPU is present, the calls raise --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfla_hip.so")

MFLA_OK = 0
MFLA_ERR_ARGUMENT = -1
MFLA_ERR_DIMENSION = -2
MFLA_ERR_ALIAS = -3
MFLA_ERR_HIP = -4
MFLA_ERR_FAILED_FACTOR = -5
MFLA_ERR_WORKSPACE = -6
MFLA_ERR_BOUNDS = -7
MFLA_NOT_CONVERGED = 1


class MFMatrix(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("ld", ctypes.c_int64),
        ("limb_stride", ctypes.c_int64),
    ]


class MFConfig(ctypes.Structure):
    _fields_ = [
        ("reduction_tile", ctypes.c_int64),
        ("column_tile", ctypes.c_int64),
        ("cholesky_block", ctypes.c_int64),
        ("lu_block", ctypes.c_int64),
        ("ldlt_block", ctypes.c_int64),
        ("ldlt_strategy", ctypes.c_int32),
        ("gemm_strategy", ctypes.c_int32),
        ("ldlt_blocked_crossover", ctypes.c_int64),
        ("thread_count", ctypes.c_int64),
        ("gemm_panel_columns", ctypes.c_int64),
        ("gemm_micro_columns", ctypes.c_int64),
        ("trsm_order", ctypes.c_int32),
        ("gemm_tile", ctypes.c_int32),
        ("gemm_packed_crossover", ctypes.c_int64),
    ]


class MFGemmPlan(ctypes.Structure):
    _fields_ = [
        ("strategy", ctypes.c_int32),
        ("reason", ctypes.c_int32),
        ("panel_columns", ctypes.c_int64),
        ("micro_columns", ctypes.c_int64),
        ("workers", ctypes.c_int64),
        ("packed_elements_per_worker", ctypes.c_int64),
        ("tile", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("tile_cols", ctypes.c_int32),
        ("network", ctypes.c_int32),
        ("launches", ctypes.c_int64),
        ("workgroups", ctypes.c_int64),
        ("remainder_rows", ctypes.c_int64),
    ]


class MFLDLTPlan(ctypes.Structure):
    _fields_ = [("strategy", ctypes.c_int32), ("reason", ctypes.c_int32), ("block_size", ctypes.c_int64)]


# Every symbol include/mfla_hip.h declares, with its ctypes prototype.
_M = MFMatrix
_I = ctypes.c_int
_I64 = ctypes.c_int64
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_CP = ctypes.POINTER(MFConfig)
PROTOTYPES = {
    "mfla_version": (_I, []),
    "mfla_last_error": (ctypes.c_char_p, []),
    "mfla_config_default": (None, [_CP]),
    "mfla_handle_create": (_I, [ctypes.POINTER(_P)]),
    "mfla_handle_destroy": (_I, [_P]),
    "mfla_handle_set_stream": (_I, [_P, _P]),
    "mfla_synchronize": (_I, [_P]),
    "mfla_workspace_create": (_I, [_P, _I, _I64, _I64, ctypes.POINTER(_P)]),
    "mfla_workspace_destroy": (_I, [_P]),
    "mfla_workspace_capacity": (_I, [_P, ctypes.POINTER(_I64)]),
    "mfla_workspace_ensure": (_I, [_P, _I64, _I64]),
    "mfla_arith": (_I, [_P, _I, _I, _I64, _P, _P, _P, _P]),
    "mfla_host_arith": (_I, [_I, _I, _I64, _DP, _DP, _DP, _DP]),
    "mfla_absmax": (_I, [_P, _I, _I, _M, _DP]),
    "mfla_all_finite": (_I, [_P, _I, _I, _M, ctypes.POINTER(_I)]),
    "mfla_copy": (_I, [_P, _I, _M, _M]),
    "mfla_from_interleaved": (_I, [_P, _I, _M, _P, _I64]),
    "mfla_to_interleaved": (_I, [_P, _I, _P, _I64, _M]),
    "mfla_gemm": (_I, [_P, _I, _M, _M, _M, _DP, _DP, _CP]),
    "mfla_gemm_plan": (_I, [_I, _I64, _I64, _I64, _CP, ctypes.POINTER(MFGemmPlan)]),
    "mfla_ldlt_plan": (_I, [_I, _I64, _CP, ctypes.POINTER(MFLDLTPlan)]),
    "mfla_syrk": (_I, [_P, _I, _I, _M, _M, _DP, _DP, _CP]),
    "mfla_gemmt": (_I, [_P, _I, _M, _M, _M, _DP, _DP, _CP]),
    "mfla_trsm": (_I, [_P, _I, _I, _I, _I, _I, _M, _M, _DP, _CP]),
    "mfla_trsv": (_I, [_P, _I, _I, _I, _I, _M, _M, _DP, _CP]),
    "mfla_gemv": (_I, [_P, _I, _I, _M, _M, _M, _DP, _DP, _CP]),
    "mfla_symv": (_I, [_P, _I, _I, _M, _M, _M, _DP, _DP, _CP]),
    "mfla_potrf": (_I, [_P, _I, _M, ctypes.POINTER(_I64), _CP]),
    "mfla_getrf": (_I, [_P, _I, _M, _P, _DP, ctypes.POINTER(_I64), _CP, _P]),
    "mfla_sytrf": (_I, [_P, _I, _M, _M, _P, _P, _DP, ctypes.POINTER(_I64), _CP, _P]),
    "mfla_sytrf_noalloc": (_I, [_P, _I, _M, _M, _P, _P, _DP, ctypes.POINTER(_I64), _CP]),
    "mfla_potrs": (_I, [_P, _I, _M, _I64, _M, _CP]),
    "mfla_getrs": (_I, [_P, _I, _M, _P, _I64, _M, _CP]),
    "mfla_sytrs": (_I, [_P, _I, _M, _M, _P, _P, _I64, _M, _CP]),
    "mfla_residual": (_I, [_P, _I, _I, _M, _M, _M, _M, _CP]),
    "mfla_residual_vector": (_I, [_P, _I, _I, _M, _M, _M, _M, _CP]),
    "mfla_residual_mixed": (_I, [_P, _I, _I, _I, _M, _M, _M, _M, _CP]),
    "mfla_add_mixed": (_I, [_P, _I, _I, _M, _M]),
    "mfla_backward_error": (_I, [_P, _I, _I, _M, _M, _M, _M, _DP]),
    "mfla_batched_solve": (_I, [_P, _I, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I, _I, _CP]),
    "mfla_batched_max_order": (_I, []),
    "mfla_batched_max_rhs": (_I, []),
    "mfla_refine_workspace_size": (_I, [_I, _I, _I64, _I64, ctypes.POINTER(_I64)]),
    "mfla_refine_solve": (_I, [_P, _I, _I, _I, _M, _M, _M, _P, _I64, ctypes.c_double, _I, ctypes.POINTER(_I64),
                               ctypes.POINTER(_I), _DP, _DP]),
    "mfla_dgemm": (_I, [_P, _I, _I, _I, _I, _M, _M, _M, ctypes.c_double, ctypes.c_double]),
    "mfla_dpotrf_inverse": (_I, [_P, _M, _M, ctypes.POINTER(_I64)]),
    "mfla_dgetrf": (_I, [_P, _M, _P, ctypes.POINTER(_I64)]),
    "mfla_geqrf": (_I, [_P, _I, _M, _M, ctypes.POINTER(_I64)]),
    "mfla_geqp3": (_I, [_P, _I, _M, _M, _P, _I, ctypes.POINTER(_I64)]),
    "mfla_ormqr": (_I, [_P, _I, _I, _M, _M, _M]),
    "mfla_trmm": (_I, [_P, _I, _I, _I, _I, _I, _M, _M, _DP, _CP]),
    "mfla_dot": (_I, [_P, _I, _M, _M, _DP]),
    "mfla_syrk_packed": (_I, [_P, _I, _M, _M, _DP, _DP, _I64, _I64, _CP]),
    "mfla_logdet": (_I, [_P, _I, _M, _I64, _DP]),
    "mfla_solve_adjoint": (_I, [_P, _I, _M, _I64, _M, _M, _M, _M, _CP]),
    "mfla_logdet_adjoint": (_I, [_P, _I, _M, _I64, _DP, _M, _CP]),
    "mfla_cholesky_adjoint": (_I, [_P, _I, _M, _I64, _M, _M, _M, _CP]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the library and set every prototype. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"libmfla_hip.so not found at {path}: build it with "
            "`python multifloatlinearalgebra.jl_amd/build.py` (hipcc, gfx950)"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().mfla_last_error()
    return msg.decode() if msg else ""
