"""The panel-fused FP64 product (mgx_dgemm_panels) against the reference's
per-panel launch sequence, through the C ABI.

The reference's SUMMA loop runs one ABMultiplyAccumulateAsync per panel
(2D-block-matrix/mmb.cu:356-358): each panel's partial product is formed
from zero and added into C.  mgx_dgemm_panels does the whole sequence in one
launch with C held in registers; it must equal the sequence of
mgx_dgemm_accumulate calls, one per panel, bit for bit (the same kernel
computes each panel's partial in the same order and the same FP64 adds join
them), and lie within the per-panel reference-order bound of the oracle.
Shapes: rank shares of 1x4, 2x2 and 4x1 grids (scaled down), ragged last
panels, panel widths that are not multiples of the 16-deep k stage, odd
widths and leading dimensions (the register-staged load path), both tiles,
overwrite and accumulate forms.
"""
import numpy as np
import pytest

from tests.gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _panels(mgx, tile, accumulate, A, B, C, m, n, k, panel, lda, ldb, ldc):
    mgx._lib.call("mgx_dgemm_panels", tile, 1 if accumulate else 0, A.data_ptr(), B.data_ptr(),
                  C.data_ptr(), m, n, k, panel, lda, ldb, ldc, 0)


def _sequence(mgx, tile, accumulate, A, B, C, m, n, k, panel, lda, ldb, ldc):
    """One accumulate launch per panel (the overwrite form zeroes C first)."""
    if not accumulate:
        C.zero_()
    for k0 in range(0, k, panel):
        kc = min(panel, k - k0)
        mgx._lib.call("mgx_dgemm_ex", tile, 1, A.data_ptr() + 8 * k0,
                      B.data_ptr() + 8 * k0 * ldb, C.data_ptr(), m, n, kc, lda, ldb, ldc, 0)


@pytest.mark.parametrize("m,n,k,panel", [
    (512, 128, 1000, 256),     # 1x4 share shape: ragged last panel of 232
    (256, 256, 1024, 256),     # 2x2 share shape: whole panels
    (128, 512, 900, 128),      # 4x1 share shape
    (300, 200, 777, 100),      # panel not a multiple of the k stage
    (193, 131, 530, 101),      # odd panel width: register-staged stages
    (64, 64, 40, 1024),        # one panel wider than k
    (1, 130, 300, 64),
])
@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("accumulate", [False, True])
def test_fused_equals_per_panel_sequence(mgx, orc, m, n, k, panel, tile, accumulate):
    import torch
    announce(f"dgemm_panels m={m} n={n} k={k} panel={panel} tile={tile} acc={accumulate}")
    A = orc.fill(m, k, 1, seed=m + 3 * k)
    B = orc.fill(k, n, 1, seed=n + 5 * k)
    C0 = orc.fill(m, n, 1, seed=7)
    dA, dB = to_dev(A), to_dev(B)
    fused = to_dev(C0)
    seq = to_dev(C0)
    _panels(mgx, tile, accumulate, dA, dB, fused, m, n, k, panel, k, n, n)
    _sequence(mgx, tile, accumulate, dA, dB, seq, m, n, k, panel, k, n, n)
    got = to_host(fused)
    assert np.array_equal(got, to_host(seq))
    # within the per-panel reference-order bound of the oracle
    ref = C0.copy() if accumulate else np.zeros((m, n))
    bound = np.zeros((m, n))
    for k0 in range(0, k, panel):
        kc = min(panel, k - k0)
        Ap = np.ascontiguousarray(A[:, k0:k0 + kc])
        Bp = np.ascontiguousarray(B[k0:k0 + kc])
        ref = orc.gemm_accumulate_order(Ap, Bp, ref, kc, 0, kc)
        absp = orc.gemm_abs_rows(Ap, Bp)
        bound += 2 * (kc * U / (1 - kc * U)) * absp + 4 * U * (np.abs(ref) + absp)
    assert np.all(np.abs(got - ref) <= bound)


def test_integer_panels_exact_and_misaligned(mgx):
    """Integer inputs are exact in any order; odd leading dimensions and an
    operand offset by one element take the register-staged path."""
    import torch
    r = np.random.default_rng(3)
    m, n, k, panel = 257, 129, 1001, 200
    announce("dgemm_panels integer, odd leading dimensions")
    lda, ldb, ldc = k + 3, n + 2, n + 5
    A = r.integers(-40, 40, size=(m, lda)).astype(np.float64)
    B = r.integers(-40, 40, size=(k + 1, ldb)).astype(np.float64)
    exact = (A[:, :k].astype(np.int64) @ B[1:, :n].astype(np.int64)).astype(np.float64)
    dA, dB = to_dev(A), to_dev(B)
    dC = torch.full((m, ldc), 5.0, dtype=torch.float64, device="cuda")
    # B starts one row in: 8-byte but not 16-byte aligned when ldb is odd
    mgx._lib.call("mgx_dgemm_panels", 1, 0, dA.data_ptr(), dB.data_ptr() + 8 * ldb,
                  dC.data_ptr(), m, n, k, panel, lda, ldb, ldc, 0)
    C = to_host(dC)
    assert np.array_equal(C[:, :n], exact)
    assert np.all(C[:, n:] == 5.0)


def test_panels_2x2_share_n16384_b256(mgx):
    """The metric share (C 8192 x 8192, K = 16384) in panels of 256: the fused
    launch equals 64 accumulate launches bit for bit, and the monolithic
    product to within FP64 reassociation."""
    import torch
    m = n = 8192
    k, panel = 16384, 256
    announce("dgemm_panels 2x2 share N=16384 b=256")
    dev = torch.device("cuda", 0)
    A = torch.empty((m, k), dtype=torch.float64, device=dev)
    B = torch.empty((k, n), dtype=torch.float64, device=dev)
    mgx.fill(A, m, k, k, 1, seed=1)
    mgx.fill(B, k, n, n, 1, seed=2)
    fused = torch.empty((m, n), dtype=torch.float64, device=dev)
    seq = torch.empty((m, n), dtype=torch.float64, device=dev)
    _panels(mgx, 1, False, A, B, fused, m, n, k, panel, k, n, n)
    _sequence(mgx, 1, False, A, B, seq, m, n, k, panel, k, n, n)
    torch.cuda.synchronize()
    assert torch.equal(fused, seq)
    rows = np.arange(0, m, 512)
    rep = mgx.error_report(A, B, fused, m, n, k, k, n, n, rows)
    assert rep["normwise"] < 1e-14
    del A, B, fused, seq
    torch.cuda.empty_cache()
