"""ctypes binding of libachr.so (the C ABI declared in include/achr.h).

The shared library is built in-tree (``achr.cu_amd/lib/libachr.so``). There is
no fallback: if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libachr.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "bin", "ACHRCuda")
WARMUP_CLI_PATH = os.path.join(os.path.dirname(_HERE), "bin", "createWarmupPts")

ACHR_OK = 0
ERROR_NAMES = {
    -1: "ACHR_E_ARG",
    -2: "ACHR_E_IO",
    -3: "ACHR_E_FORMAT",
    -4: "ACHR_E_DEVICE",
    -5: "ACHR_E_SOLVER",
    -6: "ACHR_E_STATE",
    -7: "ACHR_E_NOMEM",
}

# Every function include/achr.h declares, with its ctypes signature.
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PU32 = C.POINTER(C.c_uint32)
_PI32 = C.POINTER(C.c_int32)

SIGNATURES = {
    "achr_last_error": (C.c_char_p, []),
    "achr_version": (C.c_char_p, []),
    "achr_model_create": (_I, [_I, _I, _I, _PI, _PI, _PD, _PD, _PD, C.POINTER(_P)]),
    "achr_model_read_mps": (_I, [C.c_char_p, C.POINTER(_P)]),
    "achr_model_read_mps_rows": (_I, [C.c_char_p, C.POINTER(_P)]),
    "achr_model_file_columns": (_I, [_P, _PI]),
    "achr_model_complete_points": (_I, [_P, _PD, _I, _PD]),
    "achr_model_destroy": (None, [_P]),
    "achr_model_dims": (_I, [_P, _PI, _PI, _PI, _PI]),
    "achr_model_bounds": (_I, [_P, _PD, _PD]),
    "achr_model_csr": (_I, [_P, _PI, _PI, _PD]),
    "achr_model_compute_nullspace": (_I, [_P, _PI]),
    "achr_model_set_nullspace": (_I, [_P, _PD, _I]),
    "achr_model_get_nullspace": (_I, [_P, _PD]),
    "achr_options_default": (None, [_P]),
    "achr_sampler_create": (_I, [_P, _PD, _I, _P, C.POINTER(_P)]),
    "achr_sampler_create_csv": (_I, [_P, C.c_char_p, _P, C.POINTER(_P)]),
    "achr_sampler_destroy": (None, [_P]),
    "achr_sampler_set_stream": (_I, [_P, _P]),
    "achr_sampler_center": (_I, [_P, _PD]),
    "achr_sampler_project_warmup": (_I, [_P]),
    "achr_sampler_get_warmup": (_I, [_P, _PD, _PI]),
    "achr_sample": (_I, [_P, _I, _I, C.c_uint64, C.c_uint32, _I]),
    "achr_sample_noalloc": (_I, [_P, _I, _I, C.c_uint64, C.c_uint32, _I, _P, _P]),
    "achr_sample_points": (_I, [_P, _PD]),
    "achr_sample_points_device": (_I, [_P, C.POINTER(_P)]),
    "achr_sample_stats": (_I, [_P, _PI32]),
    "achr_sample_trace": (_I, [_P, _PD]),
    "achr_sample_timing": (_I, [_P, _PD, _PD, _PD, _PI, _PI]),
    "achr_sample_timing_check": (_I, [_P, _PD, _PI]),
    "achr_read_bandwidth": (_I, [C.c_long, _I, _PD]),
    "achr_sampler_schedule": (_I, [_P, _PI, _PI]),
    "achr_write_points_csv": (_I, [C.c_char_p, _PD, _I, _I]),
    "achr_write_points_bin": (_I, [C.c_char_p, _PD, _I, _I]),
    "achr_read_points_bin": (_I, [C.c_char_p, C.POINTER(_PD), _PI, _PI]),
    "achr_read_warmup_csv": (_I, [C.c_char_p, _I, C.POINTER(_PD), _PI]),
    "achr_free": (None, [_P]),
    "achr_philox4x32_10": (None, [_PU32, _PU32, _PU32]),
    "achr_dirnet_create": (_I, [_I, _I, _I, C.c_uint64, C.POINTER(_P)]),
    "achr_dirnet_destroy": (None, [_P]),
    "achr_dirnet_param_count": (_I, [_P, C.POINTER(C.c_long)]),
    "achr_dirnet_get_params": (_I, [_P, C.POINTER(C.c_float)]),
    "achr_dirnet_set_params": (_I, [_P, C.POINTER(C.c_float)]),
    "achr_dirnet_set_target": (_I, [_P, C.POINTER(C.c_float)]),
    "achr_dirnet_loss_grad": (_I, [_P, C.POINTER(C.c_float), _I, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "achr_dirnet_train": (_I, [_P, _I, _I, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "achr_dirnet_sample": (_I, [_P, C.POINTER(C.c_float), _I, _I, C.POINTER(C.c_float)]),
    "achr_learn_options_default": (None, [_P]),
    "achr_warmup_options_default": (None, [_P]),
    "achr_warmup_generate": (_I, [_P, _I, _P, _PD, _PD, _PI, _PD]),
    "achr_write_fva_csv": (_I, [C.c_char_p, _P, _PD]),
    "achr_sampler_learn_directions": (_I, [_P, _P, _PD]),
    "achr_sampler_set_directions": (_I, [_P, _PD, _I]),
    "achr_sampler_use_directions": (_I, [_P, _I]),
    "achr_sampler_get_directions": (_I, [_P, _PD, _PI]),
    "achr_sampler_set_run_options": (_I, [_P, _I, _I, _I]),
    "achr_batch_create": (_I, [C.POINTER(_P), _I, C.POINTER(_P)]),
    "achr_batch_destroy": (None, [_P]),
    "achr_batch_set_stream": (_I, [_P, _P]),
    "achr_batch_autotune": (_I, [_P, _PI, _I, C.c_uint64, _PD]),
    "achr_batch_lanes": (_I, [_P, _PI]),
    "achr_batch_set_lanes": (_I, [_P, _PI]),
    "achr_batch_sample": (_I, [_P, _PI, _I, C.c_uint64, C.c_uint32, _I]),
    "achr_batch_timing": (_I, [_P, _PD, _PI, _PI, _PI, _PI]),
    "achr_sampler_start_point": (_I, [_P, _PD]),
    "achr_sample_continue": (_I, [_P, _I]),
    "achr_sampler_save_state": (_I, [_P, C.c_char_p]),
    "achr_sampler_load_state": (_I, [_P, C.c_char_p]),
    "achr_quotient": (_I, [_PD, _PD, _PD, C.c_long]),
    "achr_lp_polish": (_I, [_I, _I, _PD, _PD, _PD, _PD, _PD, _PD, _I]),
    "achr_bank_reciprocal": (_I, [_PD, _PD, C.c_long]),
    "achr_gemm_sub_f64": (_I, [_I, _I, _I, _I, C.POINTER(C.c_float), C.POINTER(C.c_float), _PD, C.POINTER(C.c_int),
                               _I, C.POINTER(C.c_float)]),
    "achr_gemm_f32": (_I, [_I, _I, _I, C.POINTER(C.c_float), C.c_long, C.c_long, C.POINTER(C.c_float), C.c_long,
                           C.c_long, C.POINTER(C.c_float), C.c_long, C.c_long, C.c_float, C.c_float, _I]),
}


class AchrOptions(C.Structure):
    """Mirror of ``achr_options`` (include/achr.h)."""

    _fields_ = [
        ("dTol", C.c_double),
        ("uTol", C.c_double),
        ("maxMinTol", C.c_double),
        ("reprojTol", C.c_double),
        ("check_every", C.c_int),
        ("max_misses", C.c_int),
        ("lanes", C.c_int),
        ("reproject_mode", C.c_int),
        ("trace_chains", C.c_int),
        ("trace_thin", C.c_int),
        ("proj_prec", C.c_int),
        ("fp64_every", C.c_int),
        ("model_id", C.c_int),
        ("check_kernel", C.c_int),
    ]


class AchrWarmupOptions(C.Structure):
    """Mirror of ``achr_warmup_options`` (include/achr.h)."""

    _fields_ = [
        ("max_iters", C.c_int),
        ("check_every", C.c_int),
        ("tol", C.c_double),
        ("seed", C.c_uint64),
        ("polish", C.c_int),
    ]


class AchrLearnOptions(C.Structure):
    """Mirror of ``achr_learn_options`` (include/achr.h)."""

    _fields_ = [
        ("rank", C.c_int),
        ("width", C.c_int),
        ("iters", C.c_int),
        ("batch", C.c_int),
        ("lr", C.c_float),
        ("bank_size", C.c_int),
        ("isotropic", C.c_int),
        ("seed", C.c_uint64),
    ]


class AchrError(RuntimeError):
    def __init__(self, code: int, where: str, text: str):
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)}: {text}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Loads libachr.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libachr.so not built at {LIB_PATH}: run __graft_entry__.build()")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, where: str) -> None:
    if rc != ACHR_OK:
        raise AchrError(rc, where, lib().achr_last_error().decode())
