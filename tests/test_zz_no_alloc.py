"""mgx_split_gemm_noalloc: the split product for callers that cannot allocate,
through the C ABI, against mgx_split_gemm with a caller workspace on the same
inputs.  Two calls in a row, the second with a smaller input, must each equal
the allocating entry point bit for bit (the shared static scratch keeps the
previous call's contents, which must not leak into the next result); the
accumulate form and a non-default stream as well; sizes beyond the static
scratch are refused before any launch; and the calls leave the device's free
memory unchanged."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def _hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _inputs(mgx, torch, m, n, k, seed):
    A = torch.empty((m, k), dtype=torch.float64, device="cuda")
    B = torch.empty((k, n), dtype=torch.float64, device="cuda")
    mgx.fill(A, m, k, k, 1, seed=seed)
    mgx.fill(B, k, n, n, 1, seed=seed + 1)
    return A, B


def _noalloc(mgx, form, slices, A, B, C, m, n, k, accumulate=False, stream=0):
    mgx._lib.call("mgx_split_gemm_noalloc", form, slices, A.data_ptr(), B.data_ptr(),
                  C.data_ptr(), m, n, k, A.shape[1], B.shape[1], C.shape[1],
                  1 if accumulate else 0, stream)


def _reference(mgx, torch, form, slices, A, B, m, n, k, C0=None):
    ws_bytes = mgx.split_workspace_bytes(form, slices, m, n, k)
    ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device="cuda")
    C = C0.clone() if C0 is not None else torch.empty((m, n), dtype=torch.float64, device="cuda")
    mgx.split_gemm(A, B, C, m, n, k, A.shape[1], B.shape[1], C.shape[1], form=form,
                   slices=slices, accumulate=C0 is not None, workspace=ws,
                   workspace_bytes=ws_bytes)
    torch.cuda.synchronize()
    return C


@pytest.mark.parametrize("form_name,slices", [("mod", 15), ("int8", 8), ("bf16", 5)])
def test_noalloc_equals_allocating_entry_point(mgx, form_name, slices):
    import torch
    form = {"mod": mgx.FORM_INT8_MOD, "int8": mgx.FORM_INT8, "bf16": mgx.FORM_BF16}[form_name]
    # the second call is smaller, so stale bytes of the first stay in the scratch
    shapes = [(1000, 900, 1024), (333, 517, 700)]
    for i, (m, n, k) in enumerate(shapes):
        announce(f"noalloc {form_name} s={slices} call {i + 1} m={m} n={n} k={k}")
        assert mgx.split_workspace_bytes(form, slices, m, n, k) <= mgx._lib.NOALLOC_SCRATCH_BYTES
        A, B = _inputs(mgx, torch, m, n, k, 17 + i)
        want = _reference(mgx, torch, form, slices, A, B, m, n, k)
        C = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
        _noalloc(mgx, form, slices, A, B, C, m, n, k)
        torch.cuda.synchronize()
        assert torch.equal(C, want), (form_name, m, n, k)


def test_noalloc_accumulate_on_a_side_stream(mgx):
    import torch
    m, n, k = 640, 384, 1000
    announce("noalloc accumulate on a side stream")
    A, B = _inputs(mgx, torch, m, n, k, 23)
    C0 = torch.empty((m, n), dtype=torch.float64, device="cuda")
    mgx.fill(C0, m, n, n, 1, seed=99)
    want = _reference(mgx, torch, mgx.FORM_INT8_MOD, 16, A, B, m, n, k, C0=C0)
    s = torch.cuda.Stream()
    for _ in range(2):
        C = C0.clone()
        torch.cuda.synchronize()
        _noalloc(mgx, mgx.FORM_INT8_MOD, 16, A, B, C, m, n, k, accumulate=True,
                 stream=s.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(C, want)


def test_noalloc_refuses_sizes_beyond_the_scratch(mgx):
    import torch
    m = n = k = 2048     # 15 moduli: 180 MiB of workspace
    announce("noalloc size limit")
    assert mgx.split_workspace_bytes(mgx.FORM_INT8_MOD, 15, m, n, k) > \
        mgx._lib.NOALLOC_SCRATCH_BYTES
    A, B = _inputs(mgx, torch, m, n, k, 5)
    C = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(mgx.MgxError) as err:
        _noalloc(mgx, mgx.FORM_INT8_MOD, 15, A, B, C, m, n, k)
    assert err.value.status == mgx._lib.MGX_ERR_WORKSPACE
    assert "no-allocation scratch" in str(err.value)
    torch.cuda.synchronize()
    assert bool(torch.all(C == 0.0))          # nothing was launched


def test_noalloc_allocates_no_device_memory(mgx):
    import torch
    m, n, k = 1024, 1024, 1024
    announce("noalloc free memory")
    A, B = _inputs(mgx, torch, m, n, k, 31)
    C = torch.empty((m, n), dtype=torch.float64, device="cuda")
    # load every kernel once through the allocating path (and the no-alloc
    # path itself), so lazy runtime set-up happens before the measurement
    want = _reference(mgx, torch, mgx.FORM_INT8_MOD, 15, A, B, m, n, k)
    _noalloc(mgx, mgx.FORM_INT8_MOD, 15, A, B, C, m, n, k)
    torch.cuda.synchronize()
    before = _hip_free_bytes()
    for _ in range(3):
        _noalloc(mgx, mgx.FORM_INT8_MOD, 15, A, B, C, m, n, k)
    torch.cuda.synchronize()
    after = _hip_free_bytes()
    assert before - after <= (2 << 20), (before, after)
    assert torch.equal(C, want)
