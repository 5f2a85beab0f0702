"""ctypes binding of libhmc.so (the C-ABI in include/hmc_engine.h).

Follows the precedent of the reference's only native call site, a ctypes
binding with raw pointers (hmclab/Helpers/InterfaceMKL.py:32-35,87-89).
There is no CPU fallback: if the library is missing, importing the engine
raises ImportError, and on a machine without a GPU every compute call raises
RuntimeError from the C side.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhmc.so")

# constants mirrored from include/hmc_engine.h
OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_INJECTED = range(7)
F64, F32, BF16 = 0, 1, 2
PRIOR_NORMAL, PRIOR_LAPLACE = 1, 2
MASS_UNIT, MASS_DIAGONAL = 0, 1
INT_LEAPFROG, INT_3STAGE, INT_4STAGE = 0, 1, 2
ENGINE_AUTO, ENGINE_GEMM, ENGINE_FUSED = 0, 1, 2

PRECISIONS = {"f64": F64, "float64": F64, "f32": F32, "float32": F32, "bf16": BF16, "bfloat16": BF16}
INTEGRATORS = {"lf": INT_LEAPFROG, "3s": INT_3STAGE, "4s": INT_4STAGE}
ENGINES = {"auto": ENGINE_AUTO, "gemm": ENGINE_GEMM, "fused": ENGINE_FUSED}


class InjectedFault(RuntimeError):
    """Raised when HMC_FAULT_INJECT names the failing entry point (tests)."""


_EXC = {
    ERR_ARG: ValueError,
    ERR_STATE: AssertionError,
    ERR_HIP: RuntimeError,
    ERR_NOMEM: MemoryError,
    ERR_UNSUPPORTED: NotImplementedError,
    ERR_INJECTED: InjectedFault,
}

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)

# (name, argtypes) — every symbol declared in include/hmc_engine.h
SIGNATURES = {
    "hmc_version": [],
    "hmc_last_error": [],
    "hmc_device_count": [C.POINTER(C.c_int)],
    "hmc_problem_create": [_i64, _i32, C.POINTER(_p)],
    "hmc_problem_set_dense": [_p, _p, _i32, _i64, _dp, _dp, _i32],
    "hmc_problem_set_csr": [_p, _ip, _ip, _p, _i32, _i64, _i64, _dp, _dp, _i32],
    "hmc_problem_add_prior": [_p, _i32, _dp, _dp, C.c_double],
    "hmc_problem_set_bounds": [_p, _dp, _dp],
    "hmc_problem_set_mass": [_p, _i32, _dp],
    "hmc_problem_set_engine": [_p, _i32],
    "hmc_problem_set_option": [_p, C.c_char_p, C.c_double],
    "hmc_problem_destroy": [_p],
    "hmc_misfit_batch": [_p, _dp, _i64, _dp],
    "hmc_misfit_batch_noalloc": [_p, _dp, _i64, _dp],
    "hmc_gradient_batch": [_p, _dp, _i64, _dp],
    "hmc_chains_create": [_p, _i64, C.POINTER(_p)],
    "hmc_chains_destroy": [_p],
    "hmc_chains_set_state": [_p, _dp],
    "hmc_chains_get_state": [_p, _dp, _dp],
    "hmc_chains_set_anchor": [_p, _dp],
    "hmc_chains_trajectory_gradient": [_p, _dp, _dp],
    "hmc_chains_propose": [_p, _dp, _dp, _dp, _i32, _i32, _ip, _dp, _dp, _dp, _dp],
    "hmc_chains_run": [_p, _i64, _i32, _i32, _dp, _i32, _i32, C.c_double, C.c_double,
                       C.c_uint64, _i64, _i64, _lp],
    "hmc_chains_set_tuning_history": [_p, _i64],
    "hmc_chains_tuning_history_count": [_p, _lp],
    "hmc_chains_read_tuning_history": [_p, _i64, _i64, _dp, _dp],
    "hmc_integrator_schedule": [_i32, _i32, _dp, _dp, _i64, _lp],
    "hmc_philox4x32_10": [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "hmc_chains_set_trace": [_p, _i64, _i64],
    "hmc_chains_trace_count": [_p, _lp],
    "hmc_chains_read_trace": [_p, _i64, _i64, _dp, _dp],
    "hmc_chains_trace_device": [_p, C.POINTER(_p), C.POINTER(_dp), C.POINTER(C.c_int)],
    "hmc_chains_reset_trace": [_p],
    "hmc_chains_enable_timing": [_p, _i32],
    "hmc_chains_timing": [_p, _i32, _dp, _lp],
    "hmc_chains_sync": [_p],
    "hmc_surrogate_create": [_i64, _i64, C.c_uint64, C.POINTER(_p)],
    "hmc_surrogate_destroy": [_p],
    "hmc_surrogate_set_mode": [_p, _i32],
    "hmc_surrogate_set_normalization": [_p, _dp, _dp, C.c_double, C.c_double],
    "hmc_surrogate_param_count": [_p, _lp],
    "hmc_surrogate_get_params": [_p, C.POINTER(C.c_float)],
    "hmc_surrogate_set_params": [_p, C.POINTER(C.c_float)],
    "hmc_surrogate_eval": [_p, _dp, _i64, _dp, _dp],
    "hmc_surrogate_loss_grad": [_p, _dp, _dp, _dp, _i64, C.c_double, C.c_double, _dp,
                                C.POINTER(C.c_float)],
    "hmc_surrogate_train": [_p, _dp, _dp, _dp, _i64, _i64, _i64, C.c_double, C.c_double,
                            C.c_double, C.c_uint64, _dp],
    "hmc_chains_set_surrogate": [_p, _p],
    "hmc_chains_set_temperatures": [_p, _dp],
    "hmc_chains_exchange": [_p, _ip, _i64, _dp, _i64, _ip],
    "hmc_chains_set_exchange": [_p, _i64, _ip, _i64, _i64],
    "hmc_chains_exchange_counts": [_p, _lp],
    "hmc_set_create": [_i64, _lp, _lp, C.POINTER(_p)],
    "hmc_set_destroy": [_p],
    "hmc_set_problem_dense": [_p, _i64, C.POINTER(C.c_float), _dp, _dp, _i32],
    "hmc_set_problem_prior": [_p, _i64, _i32, _dp, _dp],
    "hmc_set_problem_mass": [_p, _i64, _i32, _dp],
    "hmc_set_problem_bounds": [_p, _i64, _dp, _dp],
    "hmc_set_problem_integrator": [_p, _i64, _i32],
    "hmc_set_set_option": [_p, C.c_char_p, C.c_double],
    "hmc_set_tile_config": [_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "hmc_set_get_option": [_p, C.c_char_p, _dp],
    "hmc_set_set_state": [_p, _dp],
    "hmc_set_get_state": [_p, _dp, _dp],
    "hmc_set_misfit": [_p, _dp, _dp],
    "hmc_set_gradient": [_p, _dp, _dp],
    "hmc_set_propose": [_p, _dp, _dp, _dp, _ip, _ip, _dp, _dp, _dp],
    "hmc_set_run": [_p, _i64, _ip, _dp, _i32, C.c_uint64, _i64, _i32, _lp],
    "hmc_set_set_trace": [_p, _i64],
    "hmc_set_trace_count": [_p, _lp],
    "hmc_set_read_trace": [_p, _i64, _i64, _dp, _dp],
    "hmc_set_enable_timing": [_p, _i32],
    "hmc_set_timing": [_p, _i32, _dp, _lp],
}

_LIB = None


def lib():
    """Load libhmc.so once; raise ImportError if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP engine with `python -m hmclab_amd.build`"
            )
        # PyTorch ships its own HIP and HSA runtimes under the same sonames as
        # the ROCm ones this library links.  Whichever set is mapped first
        # serves the whole process, and torch's device initialisation fails
        # ("No HIP GPUs are available") on the set it was not built with, so
        # when torch is installed it is imported before this library loads.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "hmc_last_error" else C.c_int
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().hmc_last_error().decode(errors="replace")
        raise _EXC.get(rc, RuntimeError)(msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def device_count() -> int:
    n = C.c_int(0)
    call("hmc_device_count", C.byref(n))
    return n.value


def dptr(a: np.ndarray | None):
    """double* of a C-contiguous float64 array (or NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)


def vptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def as_f64_rows(a, rows: int, cols: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(rows, cols))
    return a
