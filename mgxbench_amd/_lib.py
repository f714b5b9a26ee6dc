"""ctypes binding of libmgx.so (declared in include/mgx.h).

The library is loaded from the package directory; a missing library is an
error (there is no fallback implementation of any entry point).
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

# MGX_LIBRARY names an alternative build of the same library.
LIB_PATH = Path(os.environ.get("MGX_LIBRARY") or Path(__file__).resolve().parent / "libmgx.so")
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "mgx.h"

MGX_OK = 0
MGX_ERR_INVALID_ARG = 1
MGX_ERR_HIP = 2
MGX_ERR_UNSUPPORTED = 3
MGX_ERR_WORKSPACE = 4
MGX_ERR_ALLOC = 5
NOALLOC_SCRATCH_BYTES = 64 << 20   # MGX_NOALLOC_SCRATCH_BYTES

FORM_FP64 = 0
FORM_INT8 = 1
FORM_BF16 = 2
FORM_INT8_MOD = 3

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_u64 = ctypes.c_uint64

class ErrorStats(ctypes.Structure):
    _fields_ = [("normwise", ctypes.c_double), ("componentwise", ctypes.c_double),
                ("max_abs", ctypes.c_double), ("rows_checked", ctypes.c_int64)]


class SummaParams(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64),
                ("panel", ctypes.c_int64), ("form", ctypes.c_int), ("slices", ctypes.c_int),
                ("copy_mode", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("serialize", ctypes.c_int), ("split_k", ctypes.c_int64),
                ("host_threads", ctypes.c_int), ("tile", ctypes.c_int)]


class SummaStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("panels", ctypes.c_int64),
                ("host_pinned", ctypes.c_int), ("host_threads", ctypes.c_int),
                ("groups", ctypes.c_int64)]


class TuneConfig(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int), ("slices", ctypes.c_int), ("tile", ctypes.c_int),
                ("panel", ctypes.c_int64), ("split_k", ctypes.c_int64)]


TILE_DEFAULT = 0
TILE_LARGE = 1
TILE_SMALL = 2

_SIGNATURES = {
    "mgx_last_error": (ctypes.c_char_p, []),
    "mgx_version": (ctypes.c_char_p, []),
    "mgx_peak": (_dbl, [_int]),
    "mgx_dgemm": (_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _int, _vp]),
    "mgx_dgemm_accumulate": (_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                    _i64, _int, _vp]),
    "mgx_dgemm_ex": (_int, [_int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp]),
    "mgx_dgemm_panels": (_int, [_int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                                _i64, _vp]),
    "mgx_split_gemm_ex": (_int, [_int, _int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                 _i64, _int, _vp, ctypes.c_size_t, _vp]),
    "mgx_time_local_product": (_int, [ctypes.POINTER(TuneConfig), _vp, _vp, _vp, _i64, _i64, _i64,
                                      _i64, _i64, _i64, _vp, ctypes.c_size_t, _int, _int,
                                      ctypes.POINTER(_dbl)]),
    "mgx_validate_constant": (_int, [_vp, _i64, _i64, _dbl, _dbl, _dbl, _int,
                                     ctypes.POINTER(_int), ctypes.POINTER(_u64),
                                     ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]),
    "mgx_split_workspace_bytes": (ctypes.c_size_t, [_int, _int, _i64, _i64, _i64]),
    "mgx_split_gemm": (_int, [_int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                              _int, _vp, ctypes.c_size_t, _vp]),
    "mgx_split_gemm_events": (_int, [_int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                     _i64, _int, _vp, ctypes.c_size_t, _vp,
                                     ctypes.POINTER(_vp)]),
    "mgx_split_gemm_noalloc": (_int, [_int, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                      _i64, _int, _vp]),
    "mgx_split_slices_range": (_int, [_int, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "mgx_split_mod_beta": (_int, [_int, _i64]),
    "mgx_error_report": (_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _i64,
                                ctypes.POINTER(ErrorStats)]),
    "mgx_split_group_k": (_i64, [_int, _i64, _i64, _i64]),
    "mgx_summa_groups": (_i64, [_int, _i64, _i64, _i64, _vp, _i64]),
    "mgx_summa_run": (_int, [ctypes.POINTER(SummaParams), _vp, _i64, _vp, _i64, _vp, _i64,
                             ctypes.POINTER(SummaStats)]),
    "mgx_mfma_i8_ceiling": (_int, [_vp, _int, _int, _int, _vp]),
    "mgx_mfma_i8_ceiling_ops": (_dbl, [_int, _int]),
    "mgx_host_alloc": (_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "mgx_host_free": (_int, [_vp]),
    "mgx_mlp_param_count": (_i64, [_int]),
    "mgx_mlp_init": (_int, [_vp, _int, _u64]),
    "mgx_mlp_predict": (_int, [_vp, _int, _vp, _i64, _vp]),
    "mgx_mlp_gradient": (_int, [_vp, _int, _vp, _vp, _i64, _vp, _vp]),
    "mgx_mlp_train": (_int, [_vp, _int, _vp, _vp, _i64, _int, ctypes.c_float, ctypes.c_float,
                             _vp]),
    "mgx_fill": (_int, [_vp, _i64, _i64, _i64, _int, _u64, _dbl, _vp]),
    "mgx_fill_exponents": (_int, [_vp, _i64, _u64, _int, _int, _vp]),
    "mgx_scale_rows_cols": (_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp]),
}

_lib = None


class MgxError(RuntimeError):
    def __init__(self, status: int, func: str, text: str):
        super().__init__(f"{func} failed with status {status}: {text}")
        self.status = status


def header_functions() -> list[str]:
    """Names of the functions declared in include/mgx.h."""
    text = HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgx_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mgxbench_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, func: str) -> None:
    if status != MGX_OK:
        raise MgxError(status, func, lib().mgx_last_error().decode(errors="replace"))


def call(name: str, *args) -> int:
    """Calls an entry point and raises MgxError on a non-zero status."""
    status = getattr(lib(), name)(*args)
    check(status, name)
    return status
