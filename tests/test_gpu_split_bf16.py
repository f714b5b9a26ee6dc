"""Parity of the BF16 digit-slice form (MGX_FORM_BF16) through the C ABI.

The form holds the INT8 digit-slice form's digits (integers of magnitude at
most 128, exact in bf16) as bf16 planes, multiplies them on the bf16 MFMA
with FP32 sums and moves those sums into INT32 accumulators while they are
still exact integers, so its group sums, chunks and FP64 folds are the INT8
form's.  Bar: bit-exact against the oracle's model of the INT8 form
(orc_split_gemm_int8) for every slice count, including depths that split a
weight group into several INT32 chunks and span many FP32 flushes.
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce, empty, to_dev, to_host

pytestmark = pytest.mark.gpu


def _bf16(mgx, A, B, s, C0=None, tile=0):
    m, k = A.shape
    n = B.shape[1]
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(C0) if C0 is not None else empty(m, n)
    mgx._lib.call("mgx_split_gemm_ex", mgx.FORM_BF16, s, tile, dA.data_ptr(), dB.data_ptr(),
                  dC.data_ptr(), m, n, k, k, n, n, 1 if C0 is not None else 0, None, 0, 0)
    return to_host(dC)


def _same(x, y):
    return np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("s", range(1, 10))
def test_bit_exact_vs_int8_model_every_slice_count(mgx, orc, s):
    m, n, k = 130, 97, 700
    announce(f"split bf16 uniform s={s}")
    A = orc.fill(m, k, 1, seed=10 + s)
    B = orc.fill(k, n, 1, seed=20 + s)
    assert _same(_bf16(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


@pytest.mark.parametrize("m,n,k,s", [(1, 1, 1, 3), (17, 33, 5, 4), (257, 129, 63, 5),
                                     (64, 300, 1000, 7), (300, 260, 2000, 2)])
@pytest.mark.parametrize("tile", [1, 2])
def test_shapes_and_tile_argument(mgx, orc, m, n, k, s, tile):
    # the form always runs its own 256 x 128 tile; either request gives the same result
    announce(f"split bf16 shape {m}x{n}x{k} s={s} tile={tile}")
    A = orc.fill(m, k, 1, seed=m + k)
    B = orc.fill(k, n, 1, seed=n + k)
    assert _same(_bf16(mgx, A, B, s, tile=tile), orc.split_gemm_int8(A, B, s))


def test_whole_tiles_and_raster_remap(mgx, orc):
    # 1024 x 1024 is 4 x 8 whole 256 x 128 blocks: no edge checks in the fold,
    # and a block count divisible by 8 takes the XCD-aware tile remap
    m, n, k, s = 1024, 1024, 320, 3
    announce(f"split bf16 whole tiles {m}x{n}x{k} s={s}")
    A = orc.fill(m, k, 1, seed=61)
    B = orc.fill(k, n, 1, seed=62)
    assert _same(_bf16(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


@pytest.mark.parametrize("s", [3, 5, 8])
def test_stress_scaling(mgx, orc, s):
    m, n, k = 96, 80, 500
    announce(f"split bf16 stress s={s}")
    A = orc.scale_rows_cols(orc.fill(m, k, 2, seed=71), row_exp=orc.fill_exponents(m, 72, -10, 10))
    B = orc.scale_rows_cols(orc.fill(k, n, 2, seed=73), col_exp=orc.fill_exponents(n, 74, -10, 10))
    assert _same(_bf16(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


def test_several_int32_chunks_and_many_fp32_flushes(mgx, orc):
    # s = 9: the middle weight group has 9 slice pairs and an INT32 chunk of
    # 226 k tiles (14464 values), so k = 15000 takes two chunks; each chunk is
    # thousands of bf16 steps, flushed from FP32 every 16
    m, n, k, s = 48, 40, 15000, 9
    assert orc.split_chunk_tiles(s, 10) * 64 < k
    announce(f"split bf16 long k={k} s={s}")
    A = orc.fill(m, k, 1, seed=91)
    B = orc.fill(k, n, 1, seed=92)
    assert _same(_bf16(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


def test_accumulate_nonfinite_and_zero_depth(mgx, orc):
    m, n, k = 70, 60, 300
    announce("split bf16 accumulate / non-finite / k=0")
    A = orc.fill(m, k, 1, seed=81)
    B = orc.fill(k, n, 1, seed=82)
    C0 = orc.fill(m, n, 1, seed=83)
    assert _same(_bf16(mgx, A, B, 6, C0=C0), orc.split_gemm_int8(A, B, 6, C=C0))
    A[3, 5] = np.inf
    B[9, 11] = np.nan
    got = _bf16(mgx, A, B, 6)
    assert _same(got, orc.split_gemm_int8(A, B, 6))
    assert np.all(np.isnan(got[3])) and np.all(np.isnan(got[:, 11]))
    dC = empty(m, n, fill=5.0)
    mgx.split_gemm(0, 0, dC, m, n, 0, 1, n, n, form=mgx.FORM_BF16, slices=4)
    assert np.all(to_host(dC) == 0.0)


def test_workspace_and_ranges(mgx):
    lo, hi = ctypes.c_int(), ctypes.c_int()
    mgx._lib.call("mgx_split_slices_range", mgx.FORM_BF16, ctypes.byref(lo), ctypes.byref(hi))
    assert (lo.value, hi.value) == (1, 9)
    w8 = mgx.split_workspace_bytes(mgx.FORM_INT8, 5, 512, 512, 4096)
    w16 = mgx.split_workspace_bytes(mgx.FORM_BF16, 5, 512, 512, 4096)
    assert 1.9 * w8 < w16 < 2.1 * w8
    A = np.zeros((4, 10))
    with pytest.raises(mgx.MgxError):
        _bf16(mgx, A, A.T.copy(), 10)
