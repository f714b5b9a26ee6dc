"""Pins the oracle's model of the INT8 modular form (Ozaki scheme II).

The model (oracle.c orc_crt_*) scales each row of A and column of B to
integers by its own exponent (an exponent-histogram Cauchy-Schwarz rule),
forms the integer product in 128-bit arithmetic and rounds once.  It is
checked against Python's arbitrary-precision integers (a different integer
arithmetic, an independent implementation of the exponent rule and an
independently rounded conversion), the inequality the rule rests on
(sum x^2 < 4^e N / 4^12, exactly in rationals), the capacity fact the GPU's
Chinese remaindering depends on (every |C'| < P/2, on random, stress and
adversarial dense inputs), its reduction to the classical worst-case exponent
on dense lines, and an accuracy bound against the double-double product.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

EXP_BAD = 0x40000000


MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193]


def _moduli(orc, n):
    if orc is None:
        return MODULI[:n]
    got = [orc.crt_modulus(l) for l in range(n)]
    assert got == MODULI[:n]
    return got


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


W = 12          # binades resolved below a group maximum
GROUP = 16      # k values per group


def _py_line_sum(values):
    """N of one line: groups of 16 consecutive values, each nonzero finite
    value x adds 4^(12 - min(g - e(x), 12)), g = its group's largest e."""
    vals = [float(v) for v in values]
    total = 0
    for g0 in range(0, len(vals), GROUP):
        exps = [math.frexp(v)[1] for v in vals[g0:g0 + GROUP] if v != 0.0 and math.isfinite(v)]
        if exps:
            top = max(exps)
            total += sum(4 ** (W - min(top - e, W)) for e in exps)
    return total


def _py_radius2(n, k):
    P = math.prod(_moduli(None, n))
    s = math.isqrt(P // 2)
    h = math.isqrt(k)
    h += h * h < k
    R = s - (h + 1) // 2 - 1
    return R * R if R >= 1 else 0


def _py_line_beta(values, n):
    R2 = _py_radius2(n, len(values))
    N = _py_line_sum(values)
    best = -1
    for b in range(54):
        if 4 ** b * N <= R2 * 4 ** W:
            best = b
    return best


def _classical_beta(n, k):
    P = math.prod(_moduli(None, n))
    return max((b for b in range(54) if 2 * k * 4 ** b < P), default=-1)


@pytest.mark.parametrize("n,k", [(2, 1), (2, 7), (8, 1000), (14, 16384), (15, 16384),
                                 (15, 131072), (16, 16384), (16, 1), (3, 10 ** 6), (2, 40000),
                                 (15, 4096), (15, 8192), (15, 32768), (14, 1024)])
def test_dense_line_exponent_is_the_classical_one(orc, n, k):
    # a line whose elements all sit in its top binade has N = k 4^12; its
    # exponent never exceeds the classical largest beta with 2 k 4^beta < P
    # (R^2 < P/2), and equals it unless 4^beta k falls within the rounding
    # margin R^2 < P/2 - sqrt(P/2) (sqrt(k) + 4) of P/2
    P = math.prod(_moduli(orc, n))
    got = orc.crt_beta(n, k)
    want = _classical_beta(n, k)
    assert got <= want
    if want >= 0 and 4 ** want * k * 2 <= P - 2 * math.isqrt(P // 2 + 1) * (math.isqrt(k) + 5):
        assert got == want
    dense = np.full(k, 0.75)
    if k <= 40000 and got >= 0:
        assert orc.crt_line_betas(dense[None, :], n, axis=1)[0] == got


def test_exponent_of_the_metric_share(orc):
    # N = 16384, 2x2 share: k = 16384; 15 moduli give 51 for every line
    assert orc.crt_beta(15, 16384) == 51
    assert orc.crt_beta(1, 10) == -1 and orc.crt_beta(17, 10) == -1
    A = orc.fill(64, 16384, 1, seed=5)
    assert set(orc.crt_line_betas(A, 15, axis=1).tolist()) == {51}
    # uniform lines of depth 32768 gain a bit over the dense line's 50
    assert orc.crt_beta(15, 32768) == 50
    B = orc.fill(32768, 8, 1, seed=6)
    assert set(orc.crt_line_betas(B, 15, axis=0).tolist()) == {51}


def _lines(orc, seed):
    """Uniform, stress-scaled, sparse, subnormal and all-zero lines of
    several depths (not all multiples of 16)."""
    r = np.random.default_rng(seed)
    out = []
    for k in (1, 15, 17, 100, 333, 1024):
        out.append(r.uniform(-1, 1, k))
        out.append(r.choice([-1.0, 1.0], k) * r.uniform(1, 2, k) * np.ldexp(1.0, r.integers(-30, 31, k)))
        sp = np.zeros(k)
        sp[r.integers(0, k, max(1, k // 10))] = r.uniform(-3, 3, max(1, k // 10))
        out.append(sp)
        sub = r.uniform(-1, 1, k) * 2.0 ** -1070
        out.append(sub)
        out.append(np.zeros(k))
    return out


@pytest.mark.parametrize("n", [3, 9, 15, 16])
def test_line_exponents_match_an_independent_implementation(orc, n):
    for i, line in enumerate(_lines(orc, 40 + n)):
        k = len(line)
        if _py_radius2(n, k) == 0 or _classical_beta(n, k) < 0:
            continue
        want = _py_line_beta(line, n)
        assert orc.crt_line_betas(line[None, :], n, axis=1)[0] == want, (i, k)
        assert orc.crt_line_betas(line[:, None], n, axis=0)[0] == want, (i, k)
        # a line never gets less than the dense line's exponent
        assert want >= orc.crt_beta(n, k)


def test_group_sum_bounds_the_squared_norm(orc):
    # the inequality the rule rests on, in exact rationals:
    #   sum x^2 < 4^E N / 4^12,  E = frexp exponent of max |x|
    for line in _lines(orc, 7):
        nz = [float(v) for v in line if v != 0.0]
        if not nz:
            assert _py_line_sum(line) == 0
            continue
        E = math.frexp(max(abs(v) for v in nz))[1]
        lhs = sum(Fraction(v) ** 2 for v in nz)
        rhs = Fraction(2) ** (2 * E) * _py_line_sum(line) / 4 ** W
        assert lhs < rhs


@pytest.mark.parametrize("n", [2, 4, 8, 12, 14, 15, 16])
def test_products_stay_below_half_the_modulus_product(orc, n):
    # |C'_ij| < P/2 for every pair, exactly: dense same-sign lines at the top
    # of their binade (the largest C' a line pair can make), uniform and
    # stress-scaled lines, sparse lines
    P = math.prod(_moduli(orc, n))
    k = 64 if n < 6 else 700
    if orc.crt_beta(n, k) < 0:
        pytest.skip("too few moduli for this depth")
    top = 1.0 - 2.0 ** -53
    r = np.random.default_rng(n)
    rows = [np.full(k, top), np.full(k, -top), r.uniform(-1, 1, k),
            r.choice([-1.0, 1.0], k) * r.uniform(1, 2, k) * np.ldexp(1.0, r.integers(-30, 31, k))]
    sp = np.zeros(k)
    sp[:3] = [top, top, top]
    rows.append(sp)
    A = np.array(rows)
    B = np.array(rows).T.copy()
    ba = orc.crt_line_betas(A, n, axis=1)
    bb = orc.crt_line_betas(B, n, axis=0)
    e = [_exp_of_max(A[i]) for i in range(len(rows))]
    f = [_exp_of_max(B[:, j]) for j in range(len(rows))]

    def ints(values, beta, ex):
        sh = int(beta) - int(ex)
        return [int(_round_half_even(Fraction(float(v)) * Fraction(2) ** sh)) for v in values]

    Ai = [ints(A[i], ba[i], e[i]) for i in range(len(rows))]
    Bi = [ints(B[:, j], bb[j], f[j]) for j in range(len(rows))]
    worst = max(abs(sum(a * b for a, b in zip(x, y))) for x in Ai for y in Bi)
    assert 2 * worst < P
    # unless the exponent is capped at 53, the dense pair reaches a sizeable
    # share of the capacity: the bound is tight
    if orc.crt_beta(n, k) < 53:
        assert 128 * worst > P


def _round_half_even(v):
    fl = math.floor(v)
    d = v - fl
    if d > Fraction(1, 2) or (d == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return fl


def _exp_of_max(values):
    vals = [float(v) for v in values]
    if any(not math.isfinite(v) for v in vals):
        return EXP_BAD
    mx = max((abs(v) for v in vals), default=0.0)
    return 0 if mx == 0.0 else math.frexp(mx)[1]


def _py_crt(A, B, n, C=None):
    """Exact evaluation with Python integers: A'_ik = rint(A_ik 2^(beta_i - e_i))
    with beta_i from the independent implementation of the exponent rule,
    likewise B per column, C' = A'B', C = ldexp(C', e_i - beta_i + f_j - delta_j)."""
    m, k = A.shape
    nn = B.shape[1]
    e = [_exp_of_max(A[i]) for i in range(m)]
    f = [_exp_of_max(B[:, j]) for j in range(nn)]
    ba = [_py_line_beta(A[i], n) for i in range(m)]
    bb = [_py_line_beta(B[:, j], n) for j in range(nn)]

    def scaled(x, beta, ex):
        return int(_round_half_even(Fraction(float(x)) * Fraction(2) ** (beta - ex)))

    Ai = [[0 if e[i] == EXP_BAD else scaled(A[i, kk], ba[i], e[i]) for kk in range(k)]
          for i in range(m)]
    Bi = [[0 if f[j] == EXP_BAD else scaled(B[kk, j], bb[j], f[j]) for kk in range(k)]
          for j in range(nn)]
    out = np.empty((m, nn))
    for i in range(m):
        for j in range(nn):
            if e[i] == EXP_BAD or f[j] == EXP_BAD:
                x = math.nan
            else:
                cp = sum(a * b for a, b in zip(Ai[i], Bi[j]))
                # float(int) rounds to nearest
                x = math.ldexp(float(cp), e[i] - ba[i] + f[j] - bb[j])
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
    # row 0: 0.75 and two values more than 12 binades below it, so
    # N = 4^12 + 2 and the row's exponent does not depend on their exact size
    A[0] = 0.0
    A[0, 0] = 0.75
    beta = _py_line_beta(np.r_[0.75, 2.0 ** -40, 2.0 ** -40, np.zeros(k - 3)], 4)
    assert beta >= 13
    A[0, 1] = -0.5 * 2.0 ** (-beta)            # tie at the scaled unit -> 0
    A[0, 2] = 1.5 * 2.0 ** (-beta)             # tie between 1 and 2 -> 2
    assert orc.crt_line_betas(A, 4, axis=1)[0] == beta
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
    # quantisation of each line to its own bits: |dA_ik| <= 2^(e_i-beta_i-1),
    # |dB_kj| <= 2^(f_j-delta_j-1), so |C - AB|_ij <= (2^(e_i-beta_i) sum_k
    # |B_kj| + 2^(f_j-delta_j) sum_k |A_ik|)/2 + 2^(e_i-beta_i+f_j-delta_j-2) k
    # + (rounding of the result)
    m, nn, k = 24, 20, 2048
    A = orc.fill(m, k, 1, seed=41)
    B = orc.fill(k, nn, 1, seed=42)
    ba = orc.crt_line_betas(A, n, axis=1).astype(float)
    bb = orc.crt_line_betas(B, n, axis=0).astype(float)
    got = orc.crt_gemm(A, B, n)
    hi, lo = orc.gemm_dd_rows(A, B)
    e = np.array([_exp_of_max(A[i]) for i in range(m)], dtype=float) - ba
    f = np.array([_exp_of_max(B[:, j]) for j in range(nn)], dtype=float) - bb
    bound = (2.0 ** e[:, None] * np.abs(B).sum(axis=0)[None, :]
             + 2.0 ** f[None, :] * np.abs(A).sum(axis=1)[:, None]) / 2
    bound += 2.0 ** (e[:, None] + f[None, :] - 2) * k
    bound += np.abs(hi) * 2.0 ** -53
    assert np.all(np.abs(got - hi) <= bound * (1 + 1e-12))
