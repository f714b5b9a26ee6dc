"""Parity of the plain FP64-MFMA product (mgx_dgemm / mgx_dgemm_accumulate)
against the oracle, through the C ABI.

Tolerances: the MFMA kernel sums each k-block in an order the hardware fixes,
so it is not bit-identical to the reference's sequential chain.  Per element
both results lie within gamma_k * (|A||B|)_ij of the exact product, so they
differ by at most 2 * gamma_k * (|A||B|)_ij (gamma_k = k u / (1 - k u),
u = 2^-53); the test asserts exactly that bound.  Integer-valued inputs are
exact in any order and must match bit for bit.  The N=1024 config carries
the gate 1e-13 relative Frobenius error against the host product.
"""
import numpy as np
import pytest

from tests.gpu_util import announce, empty, rel_fro, to_dev, to_host

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _gamma(k):
    return k * U / (1 - k * U)


def _check_bound(C, ref, A, B, orc):
    k = A.shape[1]
    absprod = orc.gemm_abs_rows(A, B)
    diff = np.abs(C - ref)
    bound = 2 * _gamma(max(k, 1)) * absprod
    bad = diff > bound
    assert not bad.any(), f"{bad.sum()} elements beyond 2*gamma_k|A||B|, max diff {diff.max()}"


def test_known_answer_constant_inputs(mgx, orc):
    # Reference KAT: A = 1, B = 2, C = 2N exactly, validated by the reference rule.
    for n in (32, 100, 1024):
        announce(f"dgemm KAT n={n}")
        A = empty(n, n)
        B = empty(n, n)
        C = empty(n, n)
        mgx.fill(A, n, n, n, 0, value=1.0)
        mgx.fill(B, n, n, n, 0, value=2.0)
        mgx.ABMultiply(A, B, C, n, n, n, n, n, n, 32)
        assert np.all(to_host(C) == 2.0 * n)
        res = mgx.validate_matrix_C(C, n, n, 2.0 * n, 1e-9, 1e-12, rank=1, details=True)
        assert res["pass"] == 1 and res["invalid"] == 0 and res["max_abs"] == 0.0


def test_c1_uniform_n1024_gate(mgx, orc):
    n = 1024
    announce("dgemm C1 n=1024 uniform")
    A = orc.fill(n, n, 1, seed=1)
    B = orc.fill(n, n, 1, seed=2)
    dA, dB, dC = to_dev(A), to_dev(B), empty(n, n)
    mgx.ABMultiply(dA, dB, dC, n, n, n, n, n, n, 32)
    C = to_host(dC)
    hi, lo = orc.gemm_dd_rows(A, B)
    err = rel_fro(C, hi)
    assert err <= 1e-13, err
    ref = orc.gemm_kernel_order(A, B)
    assert rel_fro(ref, hi) <= 1e-13
    _check_bound(C, ref, A, B, orc)


def test_device_fill_matches_oracle_fill(mgx, orc):
    announce("fill kinds")
    for kind in (1, 2):
        x = empty(37, 53)
        mgx.fill(x, 37, 53, 53, kind, seed=99)
        assert np.array_equal(to_host(x), orc.fill(37, 53, kind, seed=99))


@pytest.mark.parametrize("m,n,k", [
    (1, 1, 1), (1, 7, 3), (17, 33, 5), (127, 129, 130), (128, 128, 16), (130, 250, 1000),
    (257, 64, 17), (64, 300, 4096), (513, 129, 31),
])
def test_ragged_shapes_uniform(mgx, orc, m, n, k):
    announce(f"dgemm ragged m={m} n={n} k={k}")
    A = orc.fill(m, k, 1, seed=m * 7 + k)
    B = orc.fill(k, n, 1, seed=n * 13 + k)
    dC = empty(m, n)
    mgx.ABMultiply(to_dev(A), to_dev(B), dC, m, n, k, k, n, n, 32)
    _check_bound(to_host(dC), orc.gemm_kernel_order(A, B), A, B, orc)


def test_integer_inputs_bit_exact(mgx, orc):
    r = np.random.default_rng(3)
    m, n, k = 200, 190, 700
    announce(f"dgemm integer m={m} n={n} k={k}")
    A = r.integers(-2000, 2000, size=(m, k)).astype(np.float64)
    B = r.integers(-2000, 2000, size=(k, n)).astype(np.float64)
    dC = empty(m, n)
    mgx.ABMultiply(to_dev(A), to_dev(B), dC, m, n, k, k, n, n, 32)
    assert np.array_equal(to_host(dC), orc.gemm_kernel_order(A, B))


def test_leading_dimensions_and_submatrices(mgx, orc):
    # Operands inside larger row-major buffers (lda > k, ldb > n, ldc > n);
    # elements outside the logical matrices must be ignored / left untouched.
    m, n, k = 70, 90, 110
    lda, ldb, ldc = 131, 97, 101
    announce(f"dgemm leading dims lda={lda} ldb={ldb} ldc={ldc}")
    Abuf = orc.fill(m, lda, 1, seed=5)
    Bbuf = orc.fill(k, ldb, 1, seed=6)
    Abuf[:, k:] = np.nan
    Bbuf[:, n:] = np.nan
    dC = empty(m, ldc, fill=-7.0)
    mgx.ABMultiply(to_dev(Abuf), to_dev(Bbuf), dC, m, n, k, lda, ldb, ldc, 32)
    C = to_host(dC)
    A = np.ascontiguousarray(Abuf[:, :k])
    B = np.ascontiguousarray(Bbuf[:, :n])
    _check_bound(C[:, :n], orc.gemm_kernel_order(A, B), A, B, orc)
    assert np.all(C[:, n:] == -7.0)


@pytest.mark.parametrize("k_offset,k_chunk", [(0, 256), (256, 256), (512, 256), (700, 256), (768, 64)])
def test_accumulate_form_matches_reference_order(mgx, orc, k_offset, k_chunk):
    # C += A[:, k_offset:k_offset+k_chunk] . B_chunk with A columns at or past
    # k_total contributing nothing (1D-row-block/mulmat_kernel.cu:100-142).
    m, n, k_total = 150, 170, 768
    announce(f"dgemm accumulate off={k_offset} chunk={k_chunk}")
    A = orc.fill(m, k_total, 1, seed=21)
    Bc = orc.fill(k_chunk, n, 1, seed=22 + k_offset)
    C0 = orc.fill(m, n, 1, seed=23)
    dC = to_dev(C0)
    mgx.ABMultiplyAccumulateAsync(to_dev(A), to_dev(Bc), dC, m, n, k_total, k_offset, k_chunk,
                                  k_total, n, n, 32)
    got = to_host(dC)
    ref = orc.gemm_accumulate_order(A, Bc, C0, k_total, k_offset, k_chunk)
    kd = max(0, min(k_chunk, k_total - k_offset))
    Ak = np.ascontiguousarray(A[:, k_offset:k_offset + kd])
    Bk = np.ascontiguousarray(Bc[:kd])
    absprod = orc.gemm_abs_rows(Ak, Bk) if kd else np.zeros((m, n))
    bound = 2 * _gamma(max(kd, 1)) * absprod + 2 * U * (np.abs(C0) + 2 * absprod)
    assert np.all(np.abs(got - ref) <= bound)


def test_accumulate_chunked_equals_whole(mgx, orc):
    # The 1D pipeline: C = 0, then one accumulate per K chunk in order
    # (1D-row-block/mmb.cu:738-858), compared with integer inputs exactly.
    r = np.random.default_rng(9)
    m, n, k, kc = 96, 80, 1000, 256
    announce("dgemm chunked pipeline integer")
    A = r.integers(-50, 50, size=(m, k)).astype(np.float64)
    B = r.integers(-50, 50, size=(k, n)).astype(np.float64)
    dA = to_dev(A)
    dC = empty(m, n, fill=0.0)
    for off in range(0, k, kc):
        mgx.ABMultiplyAccumulateAsync(dA, to_dev(B[off:off + kc]), dC, m, n, k, off, kc, k, n, n, 32)
    assert np.array_equal(to_host(dC), (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))


def test_zero_depth(mgx, orc):
    announce("dgemm zero depth")
    dC = empty(5, 6, fill=3.0)
    mgx.ABMultiply(0, 0, dC, 5, 6, 0, 1, 6, 6, 32)
    assert np.all(to_host(dC) == 0.0)
    dC = empty(5, 6, fill=3.0)
    dA = empty(5, 8, fill=1.0)
    mgx.ABMultiplyAccumulateAsync(dA, 0, dC, 5, 6, 8, 8, 4, 8, 6, 6, 32)   # k_offset == k_total
    assert np.all(to_host(dC) == 3.0)


def test_invalid_arguments_raise(mgx):
    with pytest.raises(mgx.MgxError):
        mgx.ABMultiply(0, 0, 0, 4, 4, 4, 2, 4, 4, 32)     # lda < k
    with pytest.raises(mgx.MgxError):
        mgx.ABMultiply(0, 0, 0, 4, 4, 4, 4, 4, 4, 0)      # w <= 0


def test_streams_are_honoured(mgx, orc):
    import torch
    m = n = k = 512
    announce("dgemm side stream")
    A = orc.fill(m, k, 1, seed=31)
    B = orc.fill(k, n, 1, seed=32)
    dA, dB, dC = to_dev(A), to_dev(B), empty(m, n)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        mgx.ABMultiply(dA, dB, dC, m, n, k, k, n, n, 32, stream=s)
    s.synchronize()
    _check_bound(dC.cpu().numpy(), orc.gemm_kernel_order(A, B), A, B, orc)
