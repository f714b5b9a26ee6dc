"""The library-owned split workspace (workspace == NULL) across streams and
under stream capture.

* Two NULL-workspace calls on two different streams, issued back to back
  from one thread, are ordered by the library (the second waits for the
  first's work): both results equal calls with their own workspaces, bit for
  bit.
* Inside a stream capture the scratch cannot grow: the call returns
  MGX_ERR_WORKSPACE before queueing anything, and the capture stays usable.
* A captured call that fits the existing scratch on the stream that used it
  last records normally; every replay of the graph, with eager work of the
  program in between, equals the eager result.
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu

HIP_CAPTURE_THREAD_LOCAL = 1


def _hip():
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.hipGraphInstantiate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    lib.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.hipGraphDestroy.argtypes = [ctypes.c_void_p]
    lib.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
    return lib


def _inputs(mgx, torch, m, n, k, seed):
    A = torch.empty((m, k), dtype=torch.float64, device="cuda")
    B = torch.empty((k, n), dtype=torch.float64, device="cuda")
    mgx.fill(A, m, k, k, 1, seed=seed)
    mgx.fill(B, k, n, n, 1, seed=seed + 1)
    return A, B


def test_null_workspace_calls_on_two_streams_are_ordered(mgx):
    import torch
    m, n, k = 2048, 2048, 8192
    announce("library scratch on two streams")
    A1, B1 = _inputs(mgx, torch, m, n, k, 3)
    A2, B2 = _inputs(mgx, torch, m, n, k, 5)
    ws_bytes = mgx.split_workspace_bytes(mgx.FORM_INT8_MOD, 15, m, n, k)
    ws1 = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    ws2 = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    want1 = torch.empty((m, n), dtype=torch.float64, device="cuda")
    want2 = torch.empty((m, n), dtype=torch.float64, device="cuda")
    mgx.split_gemm(A1, B1, want1, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                   workspace=ws1, workspace_bytes=ws_bytes)
    mgx.split_gemm(A2, B2, want2, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                   workspace=ws2, workspace_bytes=ws_bytes)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        C1 = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
        C2 = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        mgx.split_gemm(A1, B1, C1, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                       stream=s1)
        mgx.split_gemm(A2, B2, C2, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                       stream=s2)
        torch.cuda.synchronize()
        assert torch.equal(C1, want1) and torch.equal(C2, want2)


def _capture_case():
    """Runs in a fresh process, so the library scratch starts empty."""
    import torch
    import mgxbench_amd as mgx
    hip = _hip()
    stream = torch.cuda.Stream()
    m = n = k = 1024
    A, B = _inputs(mgx, torch, m, n, k, 9)
    C = torch.empty((m, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    # the first use of the library scratch happens inside a capture: refused
    graph = ctypes.c_void_p()
    assert hip.hipStreamBeginCapture(stream.cuda_stream, HIP_CAPTURE_THREAD_LOCAL) == 0
    try:
        mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                       stream=stream)
        raise AssertionError("a capture that must grow the scratch was accepted")
    except mgx.MgxError as err:
        assert "capture" in str(err), err
    assert hip.hipStreamEndCapture(stream.cuda_stream, ctypes.byref(graph)) == 0
    hip.hipGraphDestroy(graph)

    # one eager call on this stream sizes the scratch; a captured call of the
    # same size then records, and every replay equals the eager result
    mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15, stream=stream)
    torch.cuda.synchronize()
    eager = C.clone()
    C.fill_(float("nan"))
    torch.cuda.synchronize()
    assert hip.hipStreamBeginCapture(stream.cuda_stream, HIP_CAPTURE_THREAD_LOCAL) == 0
    mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15, stream=stream)
    assert hip.hipStreamEndCapture(stream.cuda_stream, ctypes.byref(graph)) == 0
    exe = ctypes.c_void_p()
    assert hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
    try:
        for rep in range(4):
            C.fill_(float("nan"))
            torch.cuda.synchronize()
            assert hip.hipGraphLaunch(exe, stream.cuda_stream) == 0
            torch.cuda.synchronize()
            assert torch.equal(C, eager), rep
            # eager work of the program between replays
            mgx.ABMultiply(A, B, C, m, n, k, k, n, n, stream=stream)
            torch.cuda.synchronize()
    finally:
        hip.hipGraphExecDestroy(exe)
        hip.hipGraphDestroy(graph)
    print("capture case ok", flush=True)


def test_capture_cannot_grow_the_scratch_but_can_reuse_it():
    import subprocess
    import sys
    from pathlib import Path
    announce("library scratch under stream capture (fresh process)")
    root = Path(__file__).resolve().parents[1]
    proc = subprocess.run([sys.executable, "-c",
                           "import tests.test_gpu_scratch as t; t._capture_case()"],
                          cwd=root, capture_output=True, text=True, timeout=110)
    assert proc.returncode == 0 and "capture case ok" in proc.stdout, proc.stdout + proc.stderr
