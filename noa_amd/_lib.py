"""ctypes binding of libnoa_dcs.so (the C ABI declared in include/noa_dcs.h).

The library is the only compute path: if it is missing this module raises at
load time, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnoa_dcs.so")

NOA_DCS_OK = 0
NOA_DCS_EINVAL = -1
NOA_DCS_ELAUNCH = -2
NOA_DCS_ERANGE = -3
NOA_DCS_ESTATE = -4

PREC_F64 = 0
PREC_F32MIXED = 1


class NoaDcsError(RuntimeError):
    """A C entry point returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"noa_dcs error {code}: {message}")
        self.code = code


class Element(ctypes.Structure):
    """noa_dcs_element: A (g/mol), I (GeV), Z."""
    _fields_ = [("A", ctypes.c_double), ("I", ctypes.c_double), ("Z", ctypes.c_int32)]


class OutputsF64(ctypes.Structure):
    _fields_ = [("brems", ctypes.c_void_p), ("pair", ctypes.c_void_p),
                ("photo", ctypes.c_void_p), ("ion", ctypes.c_void_p)]


class OutputsF32(ctypes.Structure):
    _fields_ = [("brems", ctypes.c_void_p), ("pair", ctypes.c_void_p),
                ("photo", ctypes.c_void_p), ("ion", ctypes.c_void_p)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("tune_cache", ctypes.c_char_p), ("unroll", ctypes.c_int32)]


class SurrogateConfig(ctypes.Structure):
    _fields_ = [("out_mean", ctypes.c_float * 4), ("out_std", ctypes.c_float * 4),
                ("trend", (ctypes.c_float * 13) * 4), ("flag_map", ctypes.c_void_p),
                ("bins", ctypes.c_int32), ("recheck_per_mille", ctypes.c_int32), ("seed", ctypes.c_uint64)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_I32 = ctypes.c_int32
_U64 = ctypes.c_uint64

# name -> (restype, argtypes); every symbol declared in include/noa_dcs.h
SIGNATURES = {
    "noa_dcs_last_error": (ctypes.c_char_p, []),
    "noa_dcs_version": (ctypes.c_char_p, []),
    "noa_dcs_eval_f64": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(Element), _D, ctypes.POINTER(OutputsF64), _P]),
    "noa_dcs_eval_f64_host": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(Element), _D,
                                             ctypes.POINTER(OutputsF64), _I32]),
    "noa_dcs_eval_f32mixed": (ctypes.c_int, [_P, _P, _I64, ctypes.POINTER(Element), _D,
                                             ctypes.POINTER(OutputsF32), _P]),
    "noa_dcs_batch_create": (ctypes.c_int, [ctypes.POINTER(Element), ctypes.POINTER(_I64), _I32, _D, _P, _P,
                                            _P, _P, _P, _P, ctypes.POINTER(BatchConfig), _P,
                                            ctypes.POINTER(_P)]),
    "noa_dcs_batch_launch": (ctypes.c_int, [_P, _P]),
    "noa_dcs_batch_info": (ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                          ctypes.POINTER(_I64)]),
    "noa_dcs_batch_destroy": (ctypes.c_int, [_P]),
    "noa_dcs_fill_workload": (ctypes.c_int, [_P, _P, _I64, _U64, _U64, _D, _D, _D, _D, _P]),
    "noa_dcs_coulomb_data_f64": (ctypes.c_int, [_P, _I64, ctypes.POINTER(Element), _D, _P, _P, _P, _P, _P]),
    "noa_dcs_coulomb_transport_f64": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _P, _P]),
    "noa_dcs_hard_scattering_f64": (ctypes.c_int, [_I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "noa_dcs_soft_scattering_workspace_bytes": (ctypes.c_int64, [_I64]),
    "noa_dcs_soft_scattering_f64": (ctypes.c_int, [_P, _I64, ctypes.POINTER(Element), _D, _P, _P, _I64, _P]),
    "noa_dcs_recoil_workspace_bytes": (ctypes.c_int64, [_I64, _I32]),
    "noa_dcs_recoil_integrals_f64": (ctypes.c_int, [_P, _I64, _D, ctypes.POINTER(Element), _D, _I32,
                                                    ctypes.POINTER(OutputsF64), ctypes.POINTER(OutputsF64), _P,
                                                    _I64, _P]),
    "noa_dcs_recoil_noalloc_scratch_bytes": (ctypes.c_int64, []),
    "noa_dcs_recoil_integrals_f64_noalloc": (ctypes.c_int, [_P, _I64, _D, ctypes.POINTER(Element), _D, _I32,
                                                            ctypes.POINTER(OutputsF64), ctypes.POINTER(OutputsF64),
                                                            _P]),
    "noa_pumas_n_table_dcs": (ctypes.c_int, [_D]),
    "noa_pumas_dcs_table": (ctypes.c_int, [_P, _P, _I32, _D, _D, _P, _I32, _P, _P, _P, _P, _P, _P]),
    "noa_pumas_dcs_f64": (ctypes.c_int, [_I32, _D, _D, _D, _P, _P, _I64, _P, _P]),
    "noa_dcs_mlp_num_params": (ctypes.c_int64, []),
    "noa_dcs_mlp_create": (ctypes.c_int, [_U64, _I32, ctypes.POINTER(_P)]),
    "noa_dcs_mlp_destroy": (ctypes.c_int, [_P]),
    "noa_dcs_mlp_get_params": (ctypes.c_int, [_P, _P, _I64]),
    "noa_dcs_mlp_set_params": (ctypes.c_int, [_P, _P, _I64]),
    "noa_dcs_mlp_features": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P]),
    "noa_dcs_mlp_forward": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "noa_dcs_mlp_forward_backward": (ctypes.c_int, [_P, _P, _P, _P, _I32, ctypes.POINTER(_D), _P, _P]),
    "noa_dcs_mlp_train": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I32, _I32, ctypes.c_float, ctypes.c_float, _U64,
                                         _P, _P]),
    "noa_dcs_mlp_infer_fused": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "noa_dcs_surrogate_eval": (ctypes.c_int, [_P, ctypes.POINTER(Element), ctypes.POINTER(_I64), _I32, _D, _P, _P,
                                              ctypes.POINTER(OutputsF32), ctypes.POINTER(SurrogateConfig), _P, _P]),
}

_lib = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Loads libnoa_dcs.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with `python -m noa_amd.build` "
                                  "(there is no CPU fallback)")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name, None)
                if fn is None:
                    continue  # reported by missing_symbols()
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def missing_symbols() -> list[str]:
    """Entry points of include/noa_dcs.h that the loaded library lacks."""
    handle = lib()
    return [n for n in SIGNATURES if getattr(handle, n, None) is None]


def last_error() -> str:
    return lib().noa_dcs_last_error().decode()


def check(rc: int) -> None:
    if rc != NOA_DCS_OK:
        raise NoaDcsError(rc, last_error())
