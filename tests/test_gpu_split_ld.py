"""Split-precision forms on submatrices: leading dimensions larger than the
operand widths (the SUMMA panel buffers and the 1D chunk pipeline pass such
views), through the C ABI.  The product of the views must equal the oracle's
model on the same values, and C outside the m x n window must not change."""
import numpy as np
import pytest

from tests.gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form_name,count", [("int8", 7), ("bf16", 4), ("mod", 15), ("mod", 9)])
@pytest.mark.parametrize("accumulate", [False, True])
def test_views_with_leading_dimensions(mgx, orc, form_name, count, accumulate):
    m, n, k = 150, 130, 333
    lda, ldb, ldc = k + 37, n + 19, n + 45
    announce(f"split {form_name} {count} views acc={accumulate}")
    r = np.random.default_rng(count + (7 if accumulate else 0))
    Abuf = r.uniform(-1, 1, size=(m + 3, lda))
    Bbuf = r.uniform(-1, 1, size=(k + 5, ldb))
    Cbuf = r.uniform(-1, 1, size=(m + 2, ldc))
    # views start one row and two columns into each buffer
    A = Abuf[1:1 + m, 2:2 + k]
    B = Bbuf[1:1 + k, 2:2 + n]
    C0 = Cbuf[1:1 + m, 2:2 + n]
    form = {"int8": mgx.FORM_INT8, "bf16": mgx.FORM_BF16, "mod": mgx.FORM_INT8_MOD}[form_name]
    dA, dB, dC = to_dev(Abuf), to_dev(Bbuf), to_dev(Cbuf)
    off = lambda t, ld: t.data_ptr() + 8 * (ld + 2)   # element (1, 2)
    mgx._lib.call("mgx_split_gemm", form, count, off(dA, lda), off(dB, ldb), off(dC, ldc), m, n, k,
                  lda, ldb, ldc, 1 if accumulate else 0, None, 0, 0)
    got = to_host(dC)
    if form_name == "mod":
        want = orc.crt_gemm(np.ascontiguousarray(A), np.ascontiguousarray(B), count,
                            C=np.ascontiguousarray(C0) if accumulate else None)
    else:
        want = orc.split_gemm_int8(np.ascontiguousarray(A), np.ascontiguousarray(B), count,
                                   C=np.ascontiguousarray(C0) if accumulate else None)
    assert np.array_equal(got[1:1 + m, 2:2 + n], want)
    outside = np.ones(Cbuf.shape, dtype=bool)
    outside[1:1 + m, 2:2 + n] = False
    assert np.array_equal(got[outside], Cbuf[outside])
