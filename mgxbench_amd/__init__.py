"""mgxbench_amd: the MGXBench local FP64 block product on MI355X (gfx950).

Python face of libmgx.  The operator functions keep the names, argument
order and meaning of the reference's operator API
(1D-row-block/mulmat_kernel.cu:86, 148, 228):

    ABMultiply(a, b, c, m, n, k, lda, ldb, ldc, w)
    ABMultiplyAccumulateAsync(a, b_chunk, c, m, n, k_total, k_offset, k_chunk,
                              lda, ldb, ldc, w, stream)
    validate_matrix_C(d_C, rows, cols, expected, abs_tol, rel_tol, rank) -> int

Matrix arguments are device pointers given as ints or as torch tensors on a
HIP device; `stream` is a torch.cuda.Stream, a raw hipStream_t int, or None
(torch's current stream).  Unlike the reference, a failing call raises
MgxError instead of aborting the process.
"""
from __future__ import annotations

from . import _lib
from ._lib import FORM_BF16, FORM_FP64, FORM_INT8, FORM_INT8_MOD, MgxError, lib

__all__ = [
    "ABMultiply", "ABMultiplyAccumulateAsync", "validate_matrix_C", "split_gemm",
    "split_gemm_events", "split_gemm_noalloc", "split_workspace_bytes", "error_report", "summa_run", "split_group_k", "summa_groups", "fill", "fill_exponents", "scale_rows_cols", "version", "peak",
    "FORM_FP64", "FORM_INT8", "FORM_INT8_MOD", "FORM_BF16", "MgxError", "lib",
]


def _ptr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


def _stream(stream) -> int:
    if stream is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def version() -> str:
    return lib().mgx_version().decode()


def peak(form: int) -> float:
    return float(lib().mgx_peak(form))


def ABMultiply(a, b, c, m, n, k, lda, ldb, ldc, w=32, stream=None) -> None:
    """C = A.B (overwrite), asynchronous on `stream` (reference: default stream)."""
    _lib.call("mgx_dgemm", _ptr(a), _ptr(b), _ptr(c), m, n, k, lda, ldb, ldc, w, _stream(stream))


def ABMultiplyAccumulateAsync(a, b_chunk, c, m, n, k_total, k_offset, k_chunk, lda, ldb, ldc,
                              w=32, stream=None) -> None:
    """C += A[:, k_offset:k_offset+k_chunk] . B_chunk, asynchronous on `stream`."""
    _lib.call("mgx_dgemm_accumulate", _ptr(a), _ptr(b_chunk), _ptr(c), m, n, k_total, k_offset,
              k_chunk, lda, ldb, ldc, w, _stream(stream))


def validate_matrix_C(d_C, rows, cols, expected, abs_tol, rel_tol, rank=0, details=False):
    """Reference rule; returns 1 (pass) or 0 (fail), or a dict when details=True."""
    import ctypes
    ok = ctypes.c_int(0)
    bad = ctypes.c_uint64(0)
    ma = ctypes.c_double(0.0)
    mr = ctypes.c_double(0.0)
    _lib.call("mgx_validate_constant", _ptr(d_C), rows, cols, expected, abs_tol, rel_tol, rank,
              ctypes.byref(ok), ctypes.byref(bad), ctypes.byref(ma), ctypes.byref(mr))
    if details:
        return {"pass": ok.value, "invalid": bad.value, "max_abs": ma.value, "max_rel": mr.value}
    return ok.value


def split_workspace_bytes(form, slices, m, n, k) -> int:
    return int(lib().mgx_split_workspace_bytes(form, slices, m, n, k))


def split_gemm(a, b, c, m, n, k, lda, ldb, ldc, *, form=FORM_INT8, slices=8, accumulate=False,
               workspace=None, workspace_bytes=0, stream=None) -> None:
    """FP64-accurate C = A.B (or C += A.B) through INT8 slice products."""
    _lib.call("mgx_split_gemm", form, slices, _ptr(a), _ptr(b), _ptr(c), m, n, k, lda, ldb, ldc,
              1 if accumulate else 0, _ptr(workspace), workspace_bytes, _stream(stream))


def split_gemm_noalloc(a, b, c, m, n, k, lda, ldb, ldc, *, form=FORM_INT8, slices=8,
                      accumulate=False, stream=None) -> None:
    """split_gemm without any device allocation: the workspace is the library's
    static per-GPU scratch (mgx_split_gemm_noalloc; sizes beyond it raise)."""
    _lib.call("mgx_split_gemm_noalloc", form, slices, _ptr(a), _ptr(b), _ptr(c), m, n, k, lda,
              ldb, ldc, 1 if accumulate else 0, _stream(stream))


def split_gemm_events(a, b, c, m, n, k, lda, ldb, ldc, events, *, form=FORM_INT8, slices=8,
                      accumulate=False, workspace=None, workspace_bytes=0, stream=None) -> None:
    """split_gemm recording 4 torch.cuda.Event objects at its phase boundaries
    (before scan, before slicing, before the product kernel, after it)."""
    import ctypes
    handles = (ctypes.c_void_p * 4)(*[int(e.cuda_event) if e is not None else 0 for e in events])
    _lib.call("mgx_split_gemm_events", form, slices, _ptr(a), _ptr(b), _ptr(c), m, n, k, lda,
              ldb, ldc, 1 if accumulate else 0, _ptr(workspace), workspace_bytes,
              _stream(stream), handles)


def error_report(a, b, c, m, n, k, lda, ldb, ldc, rows) -> dict:
    """Normwise / componentwise error of C against a GPU double-double A.B on `rows`."""
    import ctypes
    import numpy as np
    r = np.ascontiguousarray(np.asarray(rows, dtype=np.int64))
    out = _lib.ErrorStats()
    _lib.call("mgx_error_report", _ptr(a), _ptr(b), _ptr(c), m, n, k, lda, ldb, ldc,
              r.ctypes.data, len(r), ctypes.byref(out))
    return {"normwise": out.normwise, "componentwise": out.componentwise,
            "max_abs": out.max_abs, "rows_checked": out.rows_checked}


def summa_run(h_a, lda, h_b, ldb, d_c, ldc, *, m, n, k, panel, form=FORM_FP64, slices=8,
              copy_mode=0, use_graph=False, serialize=False, split_k=0,
              host_threads=0, tile=0) -> dict:
    """One rank's SUMMA panel loop: C += sum_p A_p . B_p with panels streamed
    from host strips (pointers as ints or CPU tensors; pinned ones are
    gathered by DMA directly, pageable ones packed by `host_threads` host
    threads, 0 = automatic).  Panels are multiplied in groups (summa_groups;
    split_k > 0 asks for uniform groups of that many columns) by products on
    block tile `tile` (TILE_*).  Returns the loop time, panel and group counts."""
    import ctypes
    P = _lib.SummaParams(m, n, k, panel, form, slices, copy_mode, 1 if use_graph else 0,
                         1 if serialize else 0, split_k, host_threads, tile)
    S = _lib.SummaStats()
    _lib.call("mgx_summa_run", ctypes.byref(P), _ptr(h_a), lda, _ptr(h_b), ldb, _ptr(d_c), ldc,
              ctypes.byref(S))
    return {"seconds": S.seconds, "panels": S.panels, "host_pinned": S.host_pinned,
            "host_threads": S.host_threads, "groups": S.groups}


def split_group_k(form, panel, k, split_k=0) -> int:
    """Columns of a uniform panel group (mgx_split_group_k)."""
    g = int(_lib.lib().mgx_split_group_k(form, panel, split_k, k))
    if g < 0:
        raise MgxError(-1, "mgx_split_group_k", f"invalid panel loop: form={form} "
                       f"panel={panel} k={k}")
    return g


def summa_groups(form, panel, k, split_k=0) -> list[int]:
    """Column widths of the panel loop's launches, in order (mgx_summa_groups)."""
    import ctypes
    count = int(_lib.lib().mgx_summa_groups(form, panel, split_k, k, None, 0))
    if count < 0:
        raise MgxError(-1, "mgx_summa_groups", f"invalid panel loop: form={form} "
                       f"panel={panel} k={k}")
    out = (ctypes.c_int64 * count)()
    _lib.lib().mgx_summa_groups(form, panel, split_k, k, out, count)
    return list(out)


def fill(x, rows, cols, ld, kind, seed=0, value=0.0, stream=None) -> None:
    """kind 0: constant, 1: uniform[-1,1), 2: accuracy-stress values."""
    _lib.call("mgx_fill", _ptr(x), rows, cols, ld, kind, seed, value, _stream(stream))


def fill_exponents(e, count, seed, lo, hi, stream=None) -> None:
    _lib.call("mgx_fill_exponents", _ptr(e), count, seed, lo, hi, _stream(stream))


def scale_rows_cols(x, rows, cols, ld, row_exp=None, col_exp=None, stream=None) -> None:
    _lib.call("mgx_scale_rows_cols", _ptr(x), rows, cols, ld, _ptr(row_exp), _ptr(col_exp),
              _stream(stream))
