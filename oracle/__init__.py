"""CPU oracle for the local FP64 block product -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker (or the timed CPU restatement of the
reference product).  The product path never imports it.

numpy front-end over liboracle.so (oracle/oracle.c, built by oracle/Makefile).
Every function takes and returns C-contiguous float64 numpy arrays.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_int = ctypes.c_int
_dbl = ctypes.c_double
_u64 = ctypes.c_uint64

_SIG = {
    "orc_max_threads": (_int, []),
    "orc_splitmix64_at": (_u64, [_u64, _u64]),
    "orc_fill": (None, [_vp, _i64, _i64, _i64, _int, _u64, _dbl]),
    "orc_fill_exponents": (None, [_vp, _i64, _u64, _int, _int]),
    "orc_scale_rows_cols": (None, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "orc_gemm_kernel_order": (None, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _int]),
    "orc_gemm_accumulate_order": (_int, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                         _i64, _int]),
    "orc_gemm_dd_rows": (None, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                                _int]),
    "orc_gemm_abs_rows": (None, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _int]),
    "orc_validate_constant": (_u64, [_vp, _i64, _i64, _i64, _dbl, _dbl, _dbl, _vp, _vp]),
    "orc_validate_2d": (_i64, [_vp, _i64, _dbl]),
    "orc_split_exponents": (None, [_vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp]),
    "orc_split_value": (None, [_dbl, _int, _vp]),
    "orc_split_chunk_tiles": (_int, [_int, _int]),
    "orc_split_gemm_int8": (_int, [_int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                                   _int, _int]),
    "orc_crt_modulus": (_int, [_int]),
    "orc_crt_log2_product": (_dbl, [_int]),
    "orc_crt_beta": (_int, [_int, _i64]),
    "orc_crt_line_betas": (_int, [_int, _vp, _i64, _i64, _i64, _i64, _vp]),
    "orc_crt_gemm": (_int, [_int, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _int, _int]),
    "orc_mlp_param_count": (_i64, [_int]),
    "orc_mlp_forward": (None, [_vp, _int, _vp, _i64, _vp]),
    "orc_mlp_gradient": (_dbl, [_vp, _int, _vp, _vp, _i64, _vp]),
}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def max_threads() -> int:
    return lib().orc_max_threads()


def splitmix64_at(seed: int, i: int) -> int:
    return int(lib().orc_splitmix64_at(seed, i))


def fill(rows: int, cols: int, kind: int, seed: int = 0, value: float = 0.0) -> np.ndarray:
    x = np.empty((rows, cols), dtype=np.float64)
    lib().orc_fill(_p(x), rows, cols, cols, kind, seed, value)
    return x


def fill_exponents(count: int, seed: int, lo: int, hi: int) -> np.ndarray:
    e = np.empty(count, dtype=np.int32)
    lib().orc_fill_exponents(_p(e), count, seed, lo, hi)
    return e


def scale_rows_cols(x: np.ndarray, row_exp=None, col_exp=None) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.float64).copy()
    re = None if row_exp is None else np.ascontiguousarray(row_exp, dtype=np.int32)
    ce = None if col_exp is None else np.ascontiguousarray(col_exp, dtype=np.int32)
    lib().orc_scale_rows_cols(_p(y), y.shape[0], y.shape[1], y.shape[1],
                              None if re is None else _p(re), None if ce is None else _p(ce))
    return y


def gemm_kernel_order(A: np.ndarray, B: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Reference-order product (sequential fma chain per element)."""
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    C = np.empty((m, n), dtype=np.float64)
    lib().orc_gemm_kernel_order(_p(A), _p(B), _p(C), m, n, k, max(k, 1), max(n, 1), max(n, 1),
                                nthreads)
    return C


def gemm_accumulate_order(A: np.ndarray, Bc: np.ndarray, C: np.ndarray, k_total: int,
                          k_offset: int, k_chunk: int, nthreads: int = 0) -> np.ndarray:
    """Returns C + A[:, k_offset:k_offset+k_chunk].Bc in the reference's order."""
    m, lda = A.shape
    n = C.shape[1]
    A = np.ascontiguousarray(A, dtype=np.float64)
    Bc = np.ascontiguousarray(Bc, dtype=np.float64)
    out = np.ascontiguousarray(C, dtype=np.float64).copy()
    rc = lib().orc_gemm_accumulate_order(_p(A), _p(Bc), _p(out), m, n, k_total, k_offset,
                                         k_chunk, lda, max(n, 1), max(n, 1), nthreads)
    assert rc == 0
    return out


def gemm_dd_rows(A: np.ndarray, B: np.ndarray, rows=None, nthreads: int = 0):
    """Double-double product of the selected rows: returns (hi, lo)."""
    m, k = A.shape
    n = B.shape[1]
    rows = np.arange(m, dtype=np.int64) if rows is None else np.ascontiguousarray(rows, np.int64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    hi = np.empty((len(rows), n), dtype=np.float64)
    lo = np.empty_like(hi)
    lib().orc_gemm_dd_rows(_p(A), _p(B), _p(hi), _p(lo), _p(rows), len(rows), n, k, k, n, n,
                           nthreads)
    return hi, lo


def gemm_abs_rows(A: np.ndarray, B: np.ndarray, rows=None, nthreads: int = 0) -> np.ndarray:
    m, k = A.shape
    n = B.shape[1]
    rows = np.arange(m, dtype=np.int64) if rows is None else np.ascontiguousarray(rows, np.int64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    D = np.empty((len(rows), n), dtype=np.float64)
    lib().orc_gemm_abs_rows(_p(A), _p(B), _p(D), _p(rows), len(rows), n, k, k, n, n, nthreads)
    return D


def validate_constant(C: np.ndarray, expected: float, abs_tol: float, rel_tol: float):
    C = np.ascontiguousarray(C, dtype=np.float64)
    ma = np.zeros(1)
    mr = np.zeros(1)
    bad = lib().orc_validate_constant(_p(C), C.shape[0], C.shape[1], C.shape[1], expected,
                                      abs_tol, rel_tol, _p(ma), _p(mr))
    return int(bad), float(ma[0]), float(mr[0])


def validate_2d(C: np.ndarray, expected: float) -> int:
    C = np.ascontiguousarray(C, dtype=np.float64)
    return int(lib().orc_validate_2d(_p(C), C.size, expected))


def split_exponents(A: np.ndarray, B: np.ndarray):
    m, k = A.shape
    n = B.shape[1]
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    re = np.empty(m, dtype=np.int32)
    ce = np.empty(n, dtype=np.int32)
    lib().orc_split_exponents(_p(A), _p(B), m, n, k, max(k, 1), max(n, 1), _p(re), _p(ce))
    return re, ce


def split_value(x: float, s: int) -> np.ndarray:
    out = np.zeros(s, dtype=np.int8)
    lib().orc_split_value(x, s, _p(out))
    return out


def split_chunk_tiles(s: int, t: int) -> int:
    return int(lib().orc_split_chunk_tiles(s, t))


def split_gemm_int8(A: np.ndarray, B: np.ndarray, s: int, C: np.ndarray | None = None,
                    nthreads: int = 0) -> np.ndarray:
    """Bit-exact model of libmgx's INT8 split product; C given => accumulate form."""
    m, k = A.shape
    n = B.shape[1]
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    acc = C is not None
    out = np.ascontiguousarray(C, dtype=np.float64).copy() if acc else np.empty((m, n))
    rc = lib().orc_split_gemm_int8(s, _p(A), _p(B), _p(out), m, n, k, max(k, 1), max(n, 1),
                                   max(n, 1), 1 if acc else 0, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_split_gemm_int8 failed: {rc}")
    return out


def crt_modulus(l: int) -> int:
    return int(lib().orc_crt_modulus(l))


def crt_log2_product(nmod: int) -> float:
    return float(lib().orc_crt_log2_product(nmod))


def crt_beta(nmod: int, k: int) -> int:
    """Exponent every line gets at least in the modular form (the exponent of
    a line whose elements all sit in its top binade); -1 when n moduli cannot
    hold depth k."""
    return int(lib().orc_crt_beta(nmod, k))


def crt_line_betas(X: np.ndarray, nmod: int, axis: int = 1) -> np.ndarray:
    """Per-line scaling exponents of the modular form: rows of X (axis=1, the
    rows of A) or columns of X (axis=0, the columns of B)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    rows, cols = X.shape
    lines, k = (rows, cols) if axis == 1 else (cols, rows)
    lstride, kstride = (cols, 1) if axis == 1 else (1, cols)
    out = np.empty(lines, dtype=np.int32)
    rc = lib().orc_crt_line_betas(nmod, _p(X), lines, k, lstride, kstride, _p(out))
    if rc != 0:
        raise ValueError(f"{nmod} moduli cannot hold depth {k}")
    return out


def crt_gemm(A: np.ndarray, B: np.ndarray, nmod: int, C: np.ndarray | None = None,
             nthreads: int = 0) -> np.ndarray:
    """Bit-exact model of libmgx's INT8 modular product; C given => accumulate form."""
    m, k = A.shape
    n = B.shape[1]
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    acc = C is not None
    out = np.ascontiguousarray(C, dtype=np.float64).copy() if acc else np.empty((m, n))
    rc = lib().orc_crt_gemm(nmod, _p(A), _p(B), _p(out), m, n, k, max(k, 1), max(n, 1),
                            max(n, 1), 1 if acc else 0, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_crt_gemm failed: {rc}")
    return out


def mlp_param_count(D: int) -> int:
    return int(lib().orc_mlp_param_count(D))


def mlp_forward(theta: np.ndarray, X: np.ndarray) -> np.ndarray:
    th = np.ascontiguousarray(theta, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.empty(X.shape[0])
    lib().orc_mlp_forward(_p(th), X.shape[1], _p(X), X.shape[0], _p(y))
    return y


def mlp_gradient(theta: np.ndarray, X: np.ndarray, t: np.ndarray):
    """(loss, gradient) of the mean squared error, in double precision."""
    th = np.ascontiguousarray(theta, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty(mlp_param_count(X.shape[1]))
    loss = lib().orc_mlp_gradient(_p(th), X.shape[1], _p(X), _p(t), X.shape[0], _p(g))
    return float(loss), g
