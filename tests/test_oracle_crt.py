"""Pins the oracle's model of the INT8 modular form (Ozaki scheme II).

The model (oracle.c orc_crt_*) scales rows of A and columns of B to integers,
forms the integer product in 128-bit arithmetic and rounds once.  It is
checked against Python's arbitrary-precision integers (a different integer
arithmetic and an independently rounded conversion), closed forms, the
coprimality and capacity facts the GPU's Chinese remaindering depends on,
and an accuracy bound against the double-double product.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

EXP_BAD = 0x40000000


def _moduli(orc, n):
    return [orc.crt_modulus(l) for l in range(n)]


def test_moduli_pairwise_coprime_and_int8_sized(orc):
    ms = _moduli(orc, 16)
    assert all(2 <= m <= 256 for m in ms)
    for i in range(16):
        for j in range(i + 1, 16):
            assert math.gcd(ms[i], ms[j]) == 1, (ms[i], ms[j])
    assert orc.crt_modulus(16) == -1 and orc.crt_modulus(-1) == -1
    # symmetric residues of every modulus fit int8
    assert all(-(m // 2) >= -128 and (m - 1) // 2 <= 127 for m in ms)


@pytest.mark.parametrize("n", [2, 5, 14, 15, 16])
def test_log2_product(orc, n):
    P = math.prod(_moduli(orc, n))
    assert abs(orc.crt_log2_product(n) - math.log2(P)) < 1e-9
    assert P < 2 ** 128


@pytest.mark.parametrize("n,k", [(2, 1), (2, 7), (8, 1000), (14, 16384), (15, 16384),
                                 (15, 131072), (16, 16384), (16, 1), (3, 10 ** 6), (2, 40000)])
def test_beta_is_the_largest_safe_exponent(orc, n, k):
    P = math.prod(_moduli(orc, n))
    want = -1
    for b in range(54):
        if 2 * k * 4 ** b < P:
            want = b
    assert orc.crt_beta(n, k) == want


def test_beta_of_the_metric_share(orc):
    # N = 16384, 2x2 share: k = 16384; 15 moduli give beta = 51
    assert orc.crt_beta(15, 16384) == 51
    assert orc.crt_beta(1, 10) == -1 and orc.crt_beta(17, 10) == -1


def _exp_of_max(values):
    vals = [float(v) for v in values]
    if any(not math.isfinite(v) for v in vals):
        return EXP_BAD
    mx = max((abs(v) for v in vals), default=0.0)
    return 0 if mx == 0.0 else math.frexp(mx)[1]


def _py_crt(A, B, n, k_beta=None, C=None):
    """Exact evaluation with Python integers: A' = rint(A 2^(beta-e)), C' = A'B'."""
    import oracle
    m, k = A.shape
    nn = B.shape[1]
    beta = oracle.crt_beta(n, k if k_beta is None else k_beta)
    e = [_exp_of_max(A[i]) for i in range(m)]
    f = [_exp_of_max(B[:, j]) for j in range(nn)]

    def scaled(x, ex):
        # rint(ldexp(x, beta - ex)) with ties to even, evaluated exactly
        v = Fraction(float(x)) * Fraction(2) ** (beta - ex)
        fl = math.floor(v)
        d = v - fl
        if d > Fraction(1, 2) or (d == Fraction(1, 2) and fl % 2 == 1):
            fl += 1
        return int(fl)

    Ai = [[0 if e[i] == EXP_BAD else scaled(A[i, kk], e[i]) for kk in range(k)] for i in range(m)]
    Bi = [[0 if f[j] == EXP_BAD else scaled(B[kk, j], f[j]) for kk in range(k)] for j in range(nn)]
    out = np.empty((m, nn))
    for i in range(m):
        for j in range(nn):
            if e[i] == EXP_BAD or f[j] == EXP_BAD:
                x = math.nan
            else:
                cp = sum(a * b for a, b in zip(Ai[i], Bi[j]))
                x = math.ldexp(float(cp), e[i] + f[j] - 2 * beta)  # float(int) rounds to nearest
            out[i, j] = x if C is None else C[i, j] + x
    return out


@pytest.mark.parametrize("n", [2, 6, 15, 16])
@pytest.mark.parametrize("m,nn,k", [(3, 4, 5), (6, 2, 33), (2, 3, 70)])
def test_crt_model_matches_exact_integers(orc, n, m, nn, k):
    A = orc.fill(m, k, 1, seed=11 * m + k)
    B = orc.fill(k, nn, 1, seed=13 * nn + k)
    got = orc.crt_gemm(A, B, n)
    assert np.array_equal(got, _py_crt(A, B, n))


def test_crt_model_stress_scaling_and_ties(orc):
    # wide per-row/column scaling plus values exactly halfway between scaled
    # integers (ties round to even)
    m, nn, k = 5, 4, 40
    A = orc.scale_rows_cols(orc.fill(m, k, 2, seed=3), row_exp=orc.fill_exponents(m, 4, -20, 20))
    B = orc.scale_rows_cols(orc.fill(k, nn, 2, seed=5), col_exp=orc.fill_exponents(nn, 6, -20, 20))
    beta = orc.crt_beta(4, k)
    A[0, 0] = 0.75
    A[0, 1] = -0.5 * 2.0 ** (-beta)            # tie at the scaled unit
    A[0, 2] = 1.5 * 2.0 ** (-beta)             # tie between 1 and 2 -> 2
    got = orc.crt_gemm(A, B, 4)
    assert np.array_equal(got, _py_crt(A, B, 4))


def test_crt_model_accumulate_and_nonfinite(orc):
    m, nn, k = 4, 5, 30
    A = orc.fill(m, k, 1, seed=21)
    B = orc.fill(k, nn, 1, seed=22)
    C0 = orc.fill(m, nn, 1, seed=23)
    assert np.array_equal(orc.crt_gemm(A, B, 9, C=C0), _py_crt(A, B, 9, C=C0))
    A[2, 7] = np.inf
    B[3, 1] = np.nan
    got = orc.crt_gemm(A, B, 9)
    assert np.all(np.isnan(got[2])) and np.all(np.isnan(got[:, 1]))
    ref = _py_crt(A, B, 9)
    assert np.array_equal(got, ref, equal_nan=True)


def test_crt_model_exact_on_integer_inputs(orc):
    # integer matrices are scaled without rounding, so C is the exact product
    r = np.random.default_rng(8)
    A = r.integers(-1000, 1000, size=(20, 300)).astype(np.float64)
    B = r.integers(-1000, 1000, size=(300, 15)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(orc.crt_gemm(A, B, 8), exact)


def test_crt_model_zero_depth_and_zero_rows(orc):
    A = orc.fill(3, 16, 1, seed=1)
    B = orc.fill(16, 4, 1, seed=2)
    A[1] = 0.0
    got = orc.crt_gemm(A, B, 6)
    assert np.all(got[1] == 0.0)
    Z = orc.crt_gemm(np.zeros((3, 0)), np.zeros((0, 4)), 6)
    assert np.all(Z == 0.0)


@pytest.mark.parametrize("n", [12, 15])
def test_crt_model_error_bound(orc, n):
    # quantisation of A and B to beta bits: |dA_ik| <= 2^(e_i-beta-1), so
    # |C - AB|_ij <= 2^-beta (2^e_i sum_k |B_kj| + 2^f_j sum_k |A_ik|)/2
    #              + 2^(e_i+f_j-2beta-2) k + (rounding of the result)
    m, nn, k = 24, 20, 2048
    A = orc.fill(m, k, 1, seed=41)
    B = orc.fill(k, nn, 1, seed=42)
    beta = orc.crt_beta(n, k)
    got = orc.crt_gemm(A, B, n)
    hi, lo = orc.gemm_dd_rows(A, B)
    e = np.array([_exp_of_max(A[i]) for i in range(m)], dtype=float)
    f = np.array([_exp_of_max(B[:, j]) for j in range(nn)], dtype=float)
    bound = (2.0 ** (e[:, None] - beta) * np.abs(B).sum(axis=0)[None, :]
             + 2.0 ** (f[None, :] - beta) * np.abs(A).sum(axis=1)[:, None]) / 2
    bound += 2.0 ** (e[:, None] + f[None, :] - 2 * beta - 2) * k
    bound += np.abs(hi) * 2.0 ** -53
    assert np.all(np.abs(got - hi) <= bound * (1 + 1e-12))
