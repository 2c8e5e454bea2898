"""ctypes binding of libcuvegas_amd.so (the C ABI declared in include/cuvegas.h).

The library is built in-tree by cuvegas_amd/build.py.  There is no fallback:
if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuvegas_amd.so")

CVG_OK = 0
CVG_ERR_INVALID = -1
CVG_ERR_COMPILE = -2
CVG_ERR_HIP = -3
CVG_ERR_ALLOC = -4
CVG_ERR_CAPACITY = -5
CVG_ERR_NODEVICE = -6

CVG_FP64 = 0
CVG_FP32C = 1


class cvg_params(C.Structure):
    _fields_ = [
        ("n_dim", C.c_int),
        ("n_intervals", C.c_int),
        ("n_strat", C.c_int),
        ("max_it", C.c_int),
        ("skip", C.c_int),
        ("precision", C.c_int),
        ("n_eval_it", C.c_longlong),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("seed", C.c_uint),
        ("stream_id", C.c_uint),
        ("run_len", C.c_int),
        ("has_domain", C.c_int),
        ("domain", C.POINTER(C.c_double)),
        ("n_theta", C.c_int),
    ]


class cvg_result(C.Structure):
    _fields_ = [
        ("result", C.c_double),
        ("error", C.c_double),
        ("chi2_dof", C.c_double),
        ("tot_nevals", C.c_longlong),
        ("loop_ms", C.c_double),
        ("compile_ms", C.c_double),
        ("setup_ms", C.c_double),
        ("iterations", C.c_int),
        ("n_cubes", C.c_int),
        ("n_strat", C.c_int),
        ("n_intervals", C.c_int),
        ("error_unbiased", C.c_double),
        ("chi2_dof_unbiased", C.c_double),
    ]


class cvg_ref_result(C.Structure):
    _fields_ = [
        ("res", C.c_double),
        ("err", C.c_double),
        ("n_strat", C.c_int),
        ("n_cubes", C.c_int),
        ("n_intervals", C.c_int),
        ("tot_nevals", C.c_longlong),
        ("tot_time", C.c_double),
        ("avg_it_time", C.c_double),
        ("seed", C.c_uint),
        ("multi_gpu", C.c_int),
        ("multi_cpu", C.c_int),
        ("gpu_num", C.c_int),
        ("omp_threads_num", C.c_int),
        ("gpu_thr", C.c_double),
    ]


class cvg_batch_item(C.Structure):
    _fields_ = [
        ("n_dim", C.c_int),
        ("n_strat", C.c_int),
        ("n_eval_it", C.c_longlong),
        ("theta_offset", C.c_int),
    ]


_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PU = C.POINTER(C.c_uint)
_PLL = C.POINTER(C.c_longlong)

# name -> (restype, argtypes); every symbol include/cuvegas.h declares.
SIGNATURES = {
    "cvg_last_error": (C.c_char_p, []),
    "cvg_device_info": (C.c_int, [_PI, _PI, C.c_char_p, C.c_int]),
    "cvg_device_rates": (C.c_int, [_PD, C.c_int]),
    "cvg_program_create": (C.c_int, [C.c_char_p, C.POINTER(cvg_params), C.c_char_p, C.POINTER(_P)]),
    "cvg_program_log": (C.c_char_p, [_P]),
    "cvg_program_code_size": (C.c_longlong, [_P]),
    "cvg_program_get_code": (C.c_longlong, [_P, _P, C.c_longlong]),
    "cvg_program_destroy": (None, [_P]),
    "cvg_state_create": (C.c_int, [_P, C.POINTER(cvg_params), C.POINTER(_P)]),
    "cvg_state_destroy": (None, [_P]),
    "cvg_state_set_stream": (C.c_int, [_P, _P]),
    "cvg_state_set_theta": (C.c_int, [_P, _PD, C.c_int]),
    "cvg_state_run": (C.c_int, [_P, C.c_int, C.c_int]),
    "cvg_state_result": (C.c_int, [_P, C.POINTER(cvg_result)]),
    "cvg_state_set_timing": (C.c_int, [_P, C.c_int]),
    "cvg_state_get_timing": (C.c_int, [_P, _PD, _PD, C.c_int]),
    "cvg_state_history": (C.c_int, [_P, _PD, _PD, _PLL, C.c_int]),
    "cvg_state_history_unbiased": (C.c_int, [_P, _PD, C.c_int]),
    "cvg_state_n_cubes": (C.c_longlong, [_P]),
    "cvg_state_get_map": (C.c_int, [_P, _PD]),
    "cvg_state_set_map": (C.c_int, [_P, _PD]),
    "cvg_state_get_nh": (C.c_int, [_P, _PI]),
    "cvg_state_set_nh": (C.c_int, [_P, _PI]),
    "cvg_state_get_dh": (C.c_int, [_P, _PD]),
    "cvg_state_get_cube_sums": (C.c_int, [_P, _PD, _PD]),
    "cvg_state_get_bins": (C.c_int, [_P, _PD, _PU]),
    "cvg_state_set_bins": (C.c_int, [_P, _PD, _PU]),
    "cvg_state_set_iteration": (C.c_int, [_P, C.c_int]),
    "cvg_state_set_history": (C.c_int, [_P, _PD, _PD, _PLL, _PD, C.c_int]),
    "cvg_state_fill": (C.c_int, [_P]),
    "cvg_state_refine": (C.c_int, [_P]),
    "cvg_state_reallocate": (C.c_int, [_P, _PD, C.c_double]),
    "cvg_state_dump": (C.c_int, [_P, C.c_int, C.c_longlong, _PLL, _PI, _PD, _PD]),
    "cvg_state_set_record": (C.c_int, [_P, C.c_int]),
    "cvg_state_get_snapshot": (C.c_int, [_P, C.c_int, _PD, _PI]),
    "cvg_state_backward": (C.c_int, [_P, _PD, _PD, _PD, _PD, _PD]),
    "cvg_state_backward_timing": (C.c_int, [_P, _PD]),
    "cvg_state_backward_noalloc": (C.c_int, [_P, _PD, _PD, _PD, _PD, _PD]),
    "cvg_batch_program_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(cvg_params), C.c_char_p,
                                           C.POINTER(_P)]),
    "cvg_batch_create": (C.c_int, [_P, C.POINTER(cvg_params), C.POINTER(cvg_batch_item), C.c_int, _PD,
                                   C.c_int, C.POINTER(_P)]),
    "cvg_batch_destroy": (None, [_P]),
    "cvg_batch_run": (C.c_int, [_P, C.c_int, C.c_int]),
    "cvg_batch_results": (C.c_int, [_P, _PD, _PD, _PD, _PLL, _PD, _PD, _PD]),
    "cvg_batch_history": (C.c_int, [_P, C.c_int, _PD, _PD, _PLL, C.c_int]),
    "cvg_batch_set_timing": (C.c_int, [_P, C.c_int]),
    "cvg_batch_get_timing": (C.c_int, [_P, _PD, _PD, C.c_int]),
    "cvg_batch_item_info": (C.c_int, [_P, C.c_int, _PI, _PI, _PLL]),
    "cvg_batch_get_map": (C.c_int, [_P, C.c_int, _PD]),
    "cvg_batch_get_nh": (C.c_int, [_P, C.c_int, _PI]),
    "cvg_batch_get_cube_sums": (C.c_int, [_P, C.c_int, _PD, _PD]),
    "cvg_integrate_cuda": (C.c_int, [C.c_longlong, C.c_int, C.c_int, C.c_longlong, C.c_int, C.c_int,
                                     C.c_int, C.c_double, C.c_double, C.c_uint, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_char_p,
                                     C.POINTER(cvg_ref_result)]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python cuvegas_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class CuvegasError(RuntimeError):
    """Raised for any non-zero status of the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc: int) -> int:
    if rc < 0:
        raise CuvegasError(rc, lib.cvg_last_error().decode(errors="replace"))
    return rc
