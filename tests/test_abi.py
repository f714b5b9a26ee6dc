"""The C-ABI library loads and exports every function include/mgx.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import subprocess

import pytest

from mgxbench_amd import _lib


def test_header_declares_expected_entry_points():
    names = _lib.header_functions()
    for required in ("mgx_dgemm", "mgx_dgemm_accumulate", "mgx_validate_constant",
                     "mgx_split_gemm", "mgx_split_workspace_bytes", "mgx_fill"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert _lib.LIB_PATH.exists(), "libmgx.so not built"
    handle = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in _lib.header_functions() if not hasattr(handle, n)]
    assert missing == []
    # the python binding covers the whole header
    assert set(_lib.header_functions()) <= set(_lib._SIGNATURES)


def test_exports_are_unmangled_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in _lib.header_functions():
        assert n in exported


def test_host_only_queries():
    lib = _lib.lib()
    assert lib.mgx_version().decode().count(".") == 2
    assert lib.mgx_peak(_lib.FORM_FP64) == pytest.approx(78.6)
    assert lib.mgx_peak(_lib.FORM_INT8) == pytest.approx(5000.0)
    assert lib.mgx_peak(-1) == pytest.approx(8000.0)
    assert lib.mgx_peak(99) < 0
    lo, hi = ctypes.c_int(), ctypes.c_int()
    assert lib.mgx_split_slices_range(_lib.FORM_INT8, ctypes.byref(lo), ctypes.byref(hi)) == 0
    assert (lo.value, hi.value) == (1, 9)
    # workspace sizing is pure host arithmetic
    ws = lib.mgx_split_workspace_bytes(_lib.FORM_INT8, 8, 8192, 8192, 16384)
    assert ws >= 8 * (8192 + 8192) * 16384
    assert lib.mgx_split_workspace_bytes(_lib.FORM_INT8, 0, 16, 16, 16) == 0


def test_argument_errors_are_reported_without_a_device():
    # Invalid arguments are rejected before any HIP call is made.
    lib = _lib.lib()
    st = lib.mgx_dgemm(None, None, None, 4, 4, 4, 4, 4, 4, 0, None)  # w = 0
    assert st == _lib.MGX_ERR_INVALID_ARG
    assert b"tile width" in lib.mgx_last_error()
    st = lib.mgx_dgemm(None, None, None, -1, 4, 4, 4, 4, 4, 32, None)
    assert st == _lib.MGX_ERR_INVALID_ARG
    st = lib.mgx_split_gemm(_lib.FORM_INT8, 12, None, None, None, 4, 4, 4, 4, 4, 4, 0, None, 0,
                            None)
    assert st == _lib.MGX_ERR_UNSUPPORTED
    # empty problems are a successful no-op
    assert lib.mgx_dgemm(None, None, None, 0, 4, 4, 4, 4, 4, 32, None) == 0


def test_mfma_ceiling_probe_host_side():
    lib = _lib.lib()
    # 8 waves x 32 MFMAs x 16 x 16 x 64 multiply-adds x 2 operations per round
    assert lib.mgx_mfma_i8_ceiling_ops(1, 2) == 8 * 2 * 32 * 16 * 16 * 64 * 2
    assert lib.mgx_mfma_i8_ceiling_ops(256, 12000) == 256 * 8 * 12000 * 32 * 32768
    # odd round counts, empty grids and NULL outputs are refused before any launch
    assert lib.mgx_mfma_i8_ceiling(None, 1, 2, 1, None) == _lib.MGX_ERR_INVALID_ARG
    assert lib.mgx_mfma_i8_ceiling(ctypes.c_void_p(16), 1, 3, 1, None) == _lib.MGX_ERR_INVALID_ARG
    assert lib.mgx_mfma_i8_ceiling(ctypes.c_void_p(16), 0, 2, 1, None) == _lib.MGX_ERR_INVALID_ARG


def test_missing_library_fails_loudly(tmp_path):
    # no CPU fallback: without libmgx every entry point raises
    code = ("import mgxbench_amd as mgx\n"
            "try:\n"
            "    mgx.lib()\n"
            "except RuntimeError as e:\n"
            "    print('raised', 'missing' in str(e))\n")
    env = dict(__import__("os").environ, MGX_LIBRARY=str(tmp_path / "absent" / "libmgx.so"))
    out = subprocess.run([__import__("sys").executable, "-c", code], capture_output=True,
                         text=True, env=env, cwd=str(_lib.LIB_PATH.parents[1]), timeout=120)
    assert out.stdout.strip() == "raised True", out.stdout + out.stderr


def test_split_size_limits_refused_before_any_launch():
    # m or k beyond 4194048 exceed the operand passes' grid: the call is
    # refused during argument checks (fake pointers are never dereferenced)
    lib = _lib.lib()
    p = ctypes.c_void_p(256)
    for m, k in ((5_000_000, 64), (64, 5_000_000)):
        st = lib.mgx_split_gemm(_lib.FORM_INT8_MOD, 15, p, p, p, m, 16, k, k, 16, 16, 0, None, 0,
                                None)
        assert st == _lib.MGX_ERR_INVALID_ARG, (m, k)
        assert b"too large" in lib.mgx_last_error()
