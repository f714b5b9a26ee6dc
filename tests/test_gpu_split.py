"""Parity of the INT8 split-precision product (mgx_split_gemm) through the C ABI.

Bar: bit-exact against the oracle's model of the split product (integer
slice sums are exact and the FP64 reconstruction is one fixed sequence of
ldexp / add), plus accuracy against the double-double oracle: at s = 7 the
normwise error must be no worse than the plain FP64 kernel's on the same
inputs.
"""
import numpy as np
import pytest

from tests.gpu_util import announce, empty, rel_fro, to_dev, to_host

pytestmark = pytest.mark.gpu


def _split(mgx, A, B, s, C0=None, k=None):
    m, kk = A.shape
    n = B.shape[1]
    dC = to_dev(C0) if C0 is not None else empty(m, n)
    mgx.split_gemm(to_dev(A), to_dev(B), dC, m, n, kk, kk, n, n, slices=s,
                   accumulate=C0 is not None)
    return to_host(dC)


def _same(x, y):
    return np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("m,n,k,s", [
    (1, 1, 1, 1), (16, 16, 64, 8), (17, 33, 5, 3), (130, 127, 129, 8), (128, 128, 256, 9),
    (200, 96, 1000, 6), (64, 300, 700, 8), (257, 129, 63, 2),
])
def test_bit_exact_vs_model_uniform(mgx, orc, m, n, k, s):
    announce(f"split uniform m={m} n={n} k={k} s={s}")
    A = orc.fill(m, k, 1, seed=3 * m + k)
    B = orc.fill(k, n, 1, seed=5 * n + k)
    assert _same(_split(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


@pytest.mark.parametrize("s", [3, 5, 7, 8, 9])
def test_bit_exact_vs_model_stress(mgx, orc, s):
    # Accuracy-stress inputs: +-U[1,2) * 2^U(-30,30), rows of A and columns of B
    # further scaled by 2^U(-10,10).
    m, n, k = 96, 80, 500
    announce(f"split stress s={s}")
    A = orc.scale_rows_cols(orc.fill(m, k, 2, seed=71), row_exp=orc.fill_exponents(m, 72, -10, 10))
    B = orc.scale_rows_cols(orc.fill(k, n, 2, seed=73), col_exp=orc.fill_exponents(n, 74, -10, 10))
    assert _same(_split(mgx, A, B, s), orc.split_gemm_int8(A, B, s))


def test_accumulate_form(mgx, orc):
    m, n, k = 150, 140, 300
    announce("split accumulate")
    A = orc.fill(m, k, 1, seed=81)
    B = orc.fill(k, n, 1, seed=82)
    C0 = orc.fill(m, n, 1, seed=83)
    assert _same(_split(mgx, A, B, 8, C0=C0), orc.split_gemm_int8(A, B, 8, C=C0))


def test_exact_on_7bit_integers(mgx, orc):
    r = np.random.default_rng(4)
    m, n, k = 100, 90, 2000
    announce("split 7-bit integers")
    A = r.integers(-127, 128, size=(m, k)).astype(np.float64)
    B = r.integers(-127, 128, size=(k, n)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for s in (3, 8):
        assert np.array_equal(_split(mgx, A, B, s), exact)


def test_nonfinite_rows_and_columns(mgx, orc):
    m, n, k = 40, 30, 70
    announce("split non-finite")
    A = orc.fill(m, k, 1, seed=91)
    B = orc.fill(k, n, 1, seed=92)
    A[3, 5] = np.inf
    B[9, 11] = np.nan
    got = _split(mgx, A, B, 8)
    assert _same(got, orc.split_gemm_int8(A, B, 8))
    assert np.all(np.isnan(got[3])) and np.all(np.isnan(got[:, 11]))


def test_zero_rows_and_depth(mgx, orc):
    m, n, k = 20, 20, 50
    announce("split zero rows / k=0")
    A = orc.fill(m, k, 1, seed=93)
    B = orc.fill(k, n, 1, seed=94)
    A[4] = 0.0
    B[:, 7] = 0.0
    assert _same(_split(mgx, A, B, 8), orc.split_gemm_int8(A, B, 8))
    dC = empty(m, n, fill=5.0)
    mgx.split_gemm(0, 0, dC, m, n, 0, 1, n, n, slices=8)
    assert np.all(to_host(dC) == 0.0)


def test_accuracy_n1024_no_worse_than_plain(mgx, orc):
    n = 1024
    announce("split accuracy n=1024")
    A = orc.fill(n, n, 1, seed=1)
    B = orc.fill(n, n, 1, seed=2)
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(n, n)
    mgx.ABMultiply(dA, dB, dC, n, n, n, n, n, n, 32)
    plain = to_host(dC)
    mgx.split_gemm(dA, dB, dC, n, n, n, n, n, n, slices=7)
    split = to_host(dC)
    hi, lo = orc.gemm_dd_rows(A, B)
    e_plain = rel_fro(plain, hi)
    e_split = rel_fro(split, hi)
    assert e_split <= e_plain, (e_split, e_plain)
    assert e_split <= 1e-15


def test_caller_workspace(mgx, orc):
    import torch
    m, n, k = 64, 64, 128
    announce("split caller workspace")
    A = orc.fill(m, k, 1, seed=95)
    B = orc.fill(k, n, 1, seed=96)
    need = mgx.split_workspace_bytes(mgx.FORM_INT8, 8, m, n, k)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    dC = empty(m, n)
    with pytest.raises(mgx.MgxError):
        mgx.split_gemm(to_dev(A), to_dev(B), dC, m, n, k, k, n, n, slices=8, workspace=ws,
                       workspace_bytes=need - 1)
    mgx.split_gemm(to_dev(A), to_dev(B), dC, m, n, k, k, n, n, slices=8, workspace=ws,
                   workspace_bytes=need)
    assert _same(to_host(dC), orc.split_gemm_int8(A, B, 8))
