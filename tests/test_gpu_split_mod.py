"""Parity of the INT8 modular split form (Ozaki scheme II, MGX_FORM_INT8_MOD)
through the C ABI.

Bar: bit-exact against the oracle's model (orc_crt_gemm), which forms the
integer product A'B' directly in 128-bit arithmetic -- so every residue
product, the reductions between K chunks and the Chinese remaindering on the
GPU must reproduce that integer exactly, and its conversion must round like
the host's.  Accuracy: with 15 moduli at the metric's depth the normwise
error against the double-double oracle must be no worse than the plain FP64
kernel's on the same inputs.
"""
import numpy as np
import pytest

from tests.gpu_util import announce, empty, rel_fro, to_dev, to_host

pytestmark = pytest.mark.gpu


def _mod(mgx, A, B, n, C0=None, tile=0):
    m, k = A.shape
    nn = B.shape[1]
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(C0) if C0 is not None else empty(m, nn)
    mgx._lib.call("mgx_split_gemm_ex", mgx.FORM_INT8_MOD, n, tile, dA.data_ptr(), dB.data_ptr(),
                  dC.data_ptr(), m, nn, k, k, nn, nn, 1 if C0 is not None else 0, None, 0, 0)
    return to_host(dC)


def _same(x, y):
    return np.array_equal(x, y, equal_nan=True)


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("m,n,k,nmod", [
    (1, 1, 1, 2), (16, 16, 64, 15), (17, 33, 5, 6), (130, 127, 129, 15), (128, 128, 256, 16),
    (200, 96, 1000, 13), (64, 300, 700, 15), (257, 129, 63, 4), (300, 260, 2000, 15),
    (96, 150, 900, 14),
])
def test_bit_exact_vs_model_uniform(mgx, orc, tile, m, n, k, nmod):
    announce(f"split mod uniform tile={tile} m={m} n={n} k={k} moduli={nmod}")
    A = orc.fill(m, k, 1, seed=3 * m + k)
    B = orc.fill(k, n, 1, seed=5 * n + k)
    assert _same(_mod(mgx, A, B, nmod, tile=tile), orc.crt_gemm(A, B, nmod))


@pytest.mark.parametrize("nmod", [3, 8, 15, 16])
def test_bit_exact_vs_model_stress(mgx, orc, nmod):
    m, n, k = 96, 80, 500
    announce(f"split mod stress moduli={nmod}")
    A = orc.scale_rows_cols(orc.fill(m, k, 2, seed=71), row_exp=orc.fill_exponents(m, 72, -10, 10))
    B = orc.scale_rows_cols(orc.fill(k, n, 2, seed=73), col_exp=orc.fill_exponents(n, 74, -10, 10))
    assert _same(_mod(mgx, A, B, nmod), orc.crt_gemm(A, B, nmod))


@pytest.mark.parametrize("nmod", [9, 13, 15])
def test_per_line_exponents_subnormals_and_histogram_edges(mgx, orc, nmod):
    # Lines whose exponent-group sums exercise every branch of the per-line
    # rule: subnormal and mixed normal / subnormal lines, elements exactly 11,
    # 12 and 13 binades below their group's maximum (the last histogram bin
    # and the floor term), groups that are all zero, a single nonzero element,
    # and depths that leave a ragged last group; rows of A and columns of B,
    # both tiles.  The line exponents differ from line to line, so a wrong
    # exponent anywhere changes the result's bits.
    r = np.random.default_rng(nmod)
    m, n, k = 70, 66, 1000 + 7
    announce(f"split mod per-line exponents moduli={nmod}")

    def lines(count, depth):
        X = r.uniform(-1, 1, (count, depth))
        for i in range(count):
            kind = i % 7
            if kind == 1:
                X[i] *= 2.0 ** -1060                     # subnormal line
            elif kind == 2:
                X[i, ::3] *= 2.0 ** -1060                # mixed normal / subnormal
            elif kind == 3:
                g = 16 * r.integers(0, depth // 16)
                X[i, g] = 1.5
                X[i, g + 1:g + 4] = [2.0 ** -11, 2.0 ** -12, 2.0 ** -13]   # bins 11, 12 and the floor
            elif kind == 4:
                X[i, : depth // 2] = 0.0                  # all-zero groups
            elif kind == 5:
                X[i] = 0.0
                X[i, r.integers(0, depth)] = -3.0         # one nonzero element
            elif kind == 6:
                X[i] *= np.ldexp(1.0, r.integers(-30, 31, depth))
        return X

    A = lines(m, k)
    B = lines(n, k).T.copy()
    ba = orc.crt_line_betas(A, nmod, axis=1)
    bb = orc.crt_line_betas(B, nmod, axis=0)
    if orc.crt_beta(nmod, k) < 53:      # 15 moduli at this depth: every line at the cap
        assert len(set(ba.tolist())) > 1 and len(set(bb.tolist())) > 1
    want = orc.crt_gemm(A, B, nmod)
    for tile in (1, 2):
        assert _same(_mod(mgx, A, B, nmod, tile=tile), want), tile


def test_large_magnitude_products(mgx, orc):
    # all-equal operands drive C' to k 2^(2 beta - 2), near the capacity bound,
    # and negative ones exercise the signed half of the range
    m, n, k = 40, 36, 300
    announce("split mod large magnitudes")
    A = np.ones((m, k))
    B = -np.ones((k, n))
    A[::2] = -1.0
    # 2 and 9 moduli run the kernel with a run-time modulus count, 14 and 15
    # its three-limb compile-time forms, 16 the four-limb one
    for nmod in (2, 9, 14, 15, 16):
        assert _same(_mod(mgx, A, B, nmod), orc.crt_gemm(A, B, nmod)), nmod


def test_accumulate_form(mgx, orc):
    m, n, k = 150, 140, 300
    announce("split mod accumulate")
    A = orc.fill(m, k, 1, seed=81)
    B = orc.fill(k, n, 1, seed=82)
    C0 = orc.fill(m, n, 1, seed=83)
    assert _same(_mod(mgx, A, B, 15, C0=C0), orc.crt_gemm(A, B, 15, C=C0))


def test_exact_on_integer_inputs(mgx, orc):
    r = np.random.default_rng(4)
    m, n, k = 100, 90, 2000
    announce("split mod integers")
    A = r.integers(-3000, 3000, size=(m, k)).astype(np.float64)
    B = r.integers(-3000, 3000, size=(k, n)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(_mod(mgx, A, B, 15), exact)


def test_nonfinite_rows_and_columns(mgx, orc):
    m, n, k = 40, 30, 70
    announce("split mod non-finite")
    A = orc.fill(m, k, 1, seed=91)
    B = orc.fill(k, n, 1, seed=92)
    A[3, 5] = np.inf
    B[9, 11] = np.nan
    got = _mod(mgx, A, B, 15)
    assert _same(got, orc.crt_gemm(A, B, 15))
    assert np.all(np.isnan(got[3])) and np.all(np.isnan(got[:, 11]))


def test_zero_rows_and_depth(mgx, orc):
    m, n, k = 20, 20, 50
    announce("split mod zero rows / k=0")
    A = orc.fill(m, k, 1, seed=93)
    B = orc.fill(k, n, 1, seed=94)
    A[4] = 0.0
    B[:, 7] = 0.0
    assert _same(_mod(mgx, A, B, 15), orc.crt_gemm(A, B, 15))
    dC = empty(m, n, fill=5.0)
    mgx.split_gemm(0, 0, dC, m, n, 0, 1, n, n, form=mgx.FORM_INT8_MOD, slices=15)
    assert np.all(to_host(dC) == 0.0)


def test_long_depth_reduces_between_chunks(mgx, orc):
    # k beyond one INT32 chunk (2046 tiles of 64) forces the in-register
    # reduction mod m between chunks
    m, n, k = 64, 48, 2046 * 64 + 4000
    announce(f"split mod long k={k}")
    A = orc.fill(m, k, 1, seed=95)
    B = orc.fill(k, n, 1, seed=96)
    assert _same(_mod(mgx, A, B, 16), orc.crt_gemm(A, B, 16))


@pytest.mark.parametrize("k,beta", [(16384, 51), (4096, 52)])
def test_residues_at_the_top_of_the_operand_range(mgx, orc, k, beta):
    # integer operands with |A'| up to 2^beta - 1 (row and column maxima in
    # [2^(beta-1), 2^beta), so A' = A): beta = 51 is the widest depth of the
    # short residue sequence, 52 takes the general one.  Entries sit next to
    # the half-way points j m + m/2 of every odd modulus, where the quotient
    # estimate can miss by one.
    assert orc.crt_beta(15, k) == beta
    m, n = 24, 20
    announce(f"split mod operand range edge k={k} beta={beta}")
    r = np.random.default_rng(beta)
    top = 2 ** beta
    mods = np.array([orc.crt_modulus(l) for l in range(1, 15)], dtype=np.int64)

    def edge(rows, cols):
        size = rows * cols
        mm = mods[r.integers(0, len(mods), size)]
        j = (top - 1) // mm - r.integers(0, 1000, size)
        v = np.minimum(j * mm + mm // 2 + r.integers(-2, 3, size), top - 1)
        return (v * r.choice(np.array([-1, 1]), size)).reshape(rows, cols)

    Ai = edge(m, k)
    Bi = edge(k, n)
    Ai[:, 0] = top - 1          # every row / column maximum in [2^(beta-1), 2^beta)
    Bi[0, :] = -(top - 1)
    A = Ai.astype(np.float64)
    B = Bi.astype(np.float64)
    assert np.array_equal(A.astype(np.int64), Ai) and np.array_equal(B.astype(np.int64), Bi)
    # every entry sits in the top binade of its line, so every line gets the
    # dense-line exponent and A' = A, B' = B
    assert set(orc.crt_line_betas(A, 15, axis=1).tolist()) == {beta}
    assert set(orc.crt_line_betas(B, 15, axis=0).tolist()) == {beta}
    assert _same(_mod(mgx, A, B, 15), orc.crt_gemm(A, B, 15))


def test_beta_matches_model_and_errors(mgx, orc):
    import ctypes
    for nmod in (2, 8, 14, 15, 16):
        for k in (1, 64, 1000, 16384, 131072):
            assert mgx._lib.lib().mgx_split_mod_beta(nmod, k) == orc.crt_beta(nmod, k)
    # too few moduli for the depth
    A = np.zeros((4, 10 ** 5))
    B = np.zeros((10 ** 5, 4))
    with pytest.raises(mgx.MgxError):
        _mod(mgx, A[:, :40000], B[:40000], 2)
    lo, hi = ctypes.c_int(), ctypes.c_int()
    mgx._lib.call("mgx_split_slices_range", mgx.FORM_INT8_MOD, ctypes.byref(lo), ctypes.byref(hi))
    assert (lo.value, hi.value) == (2, 16)
    with pytest.raises(mgx.MgxError):
        _mod(mgx, A[:, :10], B[:10], 17)


def test_caller_workspace(mgx, orc):
    import torch
    m, n, k = 64, 64, 128
    announce("split mod caller workspace")
    A = orc.fill(m, k, 1, seed=97)
    B = orc.fill(k, n, 1, seed=98)
    need = mgx.split_workspace_bytes(mgx.FORM_INT8_MOD, 15, m, n, k)
    assert need > mgx.split_workspace_bytes(mgx.FORM_INT8, 7, m, n, k) > 0
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(m, n)
    with pytest.raises(mgx.MgxError):
        mgx.split_gemm(dA, dB, dC, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15,
                       workspace=ws, workspace_bytes=need - 1)
    mgx.split_gemm(dA, dB, dC, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15, workspace=ws,
                   workspace_bytes=need)
    assert _same(to_host(dC), orc.crt_gemm(A, B, 15))


def test_accuracy_no_worse_than_plain(mgx, orc):
    m = n = 512
    k = 16384
    announce("split mod accuracy 512x512x16384")
    A = orc.fill(m, k, 1, seed=1)
    B = orc.fill(k, n, 1, seed=2)
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(m, n)
    mgx.ABMultiply(dA, dB, dC, m, n, k, k, n, n, 32)
    plain = to_host(dC)
    mgx.split_gemm(dA, dB, dC, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15)
    split = to_host(dC)
    rows = np.arange(0, m, 8)
    hi, lo = orc.gemm_dd_rows(A, B, rows=rows)
    e_plain = rel_fro(plain[rows], hi)
    e_split = rel_fro(split[rows], hi)
    assert e_split <= e_plain, (e_split, e_plain)
    assert e_split <= 1e-15
