"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each oracle function is checked against something that does not restate it:
exact rational brute force, closed forms, published vectors, or algebraic
identities that a plausible slip (dropped term, swapped operand, wrong
index, wrong sign) would break.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from tests import fractions_ref as fr


def rng_matrix(rows, cols, seed, kind=1):
    import oracle
    return oracle.fill(rows, cols, kind, seed)


# ---- synthetic inputs ------------------------------------------------------------

def test_splitmix64_published_vectors(orc):
    # Published splitmix64 outputs (Vigna's reference generator):
    # seed 0 -> e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f;
    # seed 1234567 -> 6457827717110365317 first.
    assert orc.splitmix64_at(0, 1) == 0xE220A8397B1DCDAF
    assert orc.splitmix64_at(0, 2) == 0x6E789E6AA1B965F4
    assert orc.splitmix64_at(0, 3) == 0x06C45D188009454F
    assert orc.splitmix64_at(1234567, 1) == 6457827717110365317


def test_uniform_fill_definition_and_range(orc):
    x = orc.fill(64, 96, 1, seed=7)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert abs(x.mean()) < 0.05
    # element (r, c) uses draw r*cols+c+1
    r, c = 5, 17
    raw = orc.splitmix64_at(7, r * 96 + c + 1)
    assert x[r, c] == 2.0 * (raw >> 11) * 2.0 ** -53 - 1.0


def test_stress_fill_distribution(orc):
    x = orc.fill(128, 128, 2, seed=3)
    m, e = np.frexp(np.abs(x))
    assert np.all((m >= 0.5) & (m < 1.0))
    assert e.min() >= -29 and e.max() <= 31          # |x| in [2^-30, 2^31)
    assert 0.4 < (x < 0).mean() < 0.6


# ---- reference-order product --------------------------------------------------------

@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (3, 5, 7), (4, 6, 33), (7, 3, 64)])
def test_kernel_order_matches_exact_fma_chain(orc, m, n, k):
    A = rng_matrix(m, k, 11)
    B = rng_matrix(k, n, 12)
    C = orc.gemm_kernel_order(A, B)
    ref = np.array([[fr.chain(A[i, :], B[:, j]) for j in range(n)] for i in range(m)])
    assert np.array_equal(C, ref)


def test_kernel_order_known_answer(orc):
    # Reference KAT: A = 1, B = 2 gives C = 2N exactly (1D-row-block/mmb.cu:939-946).
    for n in (1, 31, 32, 33, 257):
        C = orc.gemm_kernel_order(np.ones((n, n)), np.full((n, n), 2.0))
        assert np.all(C == 2.0 * n)


def test_kernel_order_integer_inputs_exact(orc):
    r = np.random.default_rng(0)
    A = r.integers(-1000, 1000, size=(37, 300)).astype(np.float64)
    B = r.integers(-1000, 1000, size=(300, 41)).astype(np.float64)
    exact = A.astype(np.int64) @ B.astype(np.int64)
    assert np.array_equal(orc.gemm_kernel_order(A, B), exact.astype(np.float64))


def test_kernel_order_transpose_identity(orc):
    # (A.B)^T == B^T.A^T bit for bit: same chain, commuted fma operands.
    A = rng_matrix(33, 70, 21)
    B = rng_matrix(70, 45, 22)
    C = orc.gemm_kernel_order(A, B)
    Ct = orc.gemm_kernel_order(np.ascontiguousarray(B.T), np.ascontiguousarray(A.T))
    assert np.array_equal(C, Ct.T)


def test_kernel_order_error_bound(orc):
    # |chain - exact| <= gamma_k * sum|a||b| (sequential summation bound).
    m, n, k = 6, 5, 200
    A = rng_matrix(m, k, 31)
    B = rng_matrix(k, n, 32)
    C = orc.gemm_kernel_order(A, B)
    u = 2.0 ** -53
    gamma = k * u / (1 - k * u)
    for i in range(m):
        for j in range(n):
            exact = fr.exact_dot(A[i, :], B[:, j])
            absdot = sum(abs(A[i, t] * B[t, j]) for t in range(k))
            assert abs(float(Fraction(C[i, j]) - exact)) <= gamma * absdot


def test_accumulate_order_matches_exact(orc):
    m, n, k_total = 4, 5, 40
    A = rng_matrix(m, k_total, 41)
    C0 = rng_matrix(m, n, 42)
    for k_offset, k_chunk in [(0, 16), (16, 16), (32, 16), (36, 8), (40, 8)]:
        Bc = rng_matrix(k_chunk, n, 43 + k_offset)
        out = orc.gemm_accumulate_order(A, Bc, C0, k_total, k_offset, k_chunk)
        kd = max(0, min(k_chunk, k_total - k_offset))
        for i in range(m):
            for j in range(n):
                part = fr.chain(A[i, k_offset:k_offset + kd], Bc[:kd, j])
                assert out[i, j] == fr.add(C0[i, j], part)


# ---- double-double reference ----------------------------------------------------------

def test_dd_product_is_near_exact(orc):
    m, n, k = 5, 6, 300
    A = rng_matrix(m, k, 51)
    B = rng_matrix(k, n, 52)
    hi, lo = orc.gemm_dd_rows(A, B)
    for i in range(m):
        for j in range(n):
            exact = fr.exact_dot(A[i, :], B[:, j])
            absdot = sum(abs(A[i, t] * B[t, j]) for t in range(k))
            err = abs(float(Fraction(hi[i, j]) + Fraction(lo[i, j]) - exact))
            assert err <= 2.0 ** -100 * absdot + 2.0 ** -104 * abs(float(exact))
            assert hi[i, j] == fr.rnd(exact) or abs(hi[i, j] - fr.rnd(exact)) <= math.ulp(hi[i, j])


def test_dd_selected_rows(orc):
    A = rng_matrix(20, 64, 61)
    B = rng_matrix(64, 8, 62)
    rows = np.array([3, 0, 19])
    hi, _ = orc.gemm_dd_rows(A, B, rows)
    full, _ = orc.gemm_dd_rows(A, B)
    assert np.array_equal(hi, full[rows])


def test_abs_product(orc):
    A = rng_matrix(9, 30, 71)
    B = rng_matrix(30, 7, 72)
    D = orc.gemm_abs_rows(A, B)
    np.testing.assert_allclose(D, np.abs(A) @ np.abs(B), rtol=1e-14)


# ---- validation rules -------------------------------------------------------------------

def test_validate_constant_rule(orc):
    n = 64
    e = 2.0 * n
    C = np.full((8, 8), e)
    assert orc.validate_constant(C, e, 1e-9, 1e-12)[0] == 0
    C[1, 2] = e + 1e-3          # abs and rel both above tolerance
    C[3, 4] = e + 5e-10         # abs below 1e-9 -> valid
    C[5, 6] = np.nan            # NaN comparisons are false -> valid (reference behaviour)
    bad, ma, mr = orc.validate_constant(C, e, 1e-9, 1e-12)
    assert bad == 1
    assert ma == pytest.approx(1e-3, rel=1e-6)
    assert mr == pytest.approx(1e-3 / e, rel=1e-6)
    # expected == 0: relative error is 0, so nothing is ever invalid
    assert orc.validate_constant(np.ones((2, 2)), 0.0, 1e-9, 1e-12)[0] == 0


def test_validate_2d_rule(orc):
    e = 2.0 * 4096
    C = np.full(100, e)
    assert orc.validate_2d(C, e) == -1
    C[7] = e + 2e-9 + 1e-12 * e   # just beyond 1e-9 + 1e-12|e|
    assert orc.validate_2d(C, e) == 7
    C[7] = e
    C[9] = np.inf
    assert orc.validate_2d(C, e) == 9


# ---- split-precision model -------------------------------------------------------------

def _py_slices(x: float, s: int):
    """Exact digits: rint (half to even) of x 2^6, then of 256 * remainder,
    then carries 128 -> -128 (+1 above) from the least significant digit."""
    y = Fraction(x) * 64
    dig = []
    for _ in range(s):
        d = round(y)                       # Fraction round: half to even
        dig.append(d)
        y = (y - d) * 256
    for p in range(s - 1, 0, -1):
        if dig[p] >= 128:
            dig[p] -= 256
            dig[p - 1] += 1
    return dig


@pytest.mark.parametrize("s", [1, 3, 7, 8, 9])
def test_split_value_slices(orc, s):
    r = np.random.default_rng(s)
    special = [0.0, 0.5, -0.5, 1 - 2 ** -53, -(1 - 2 ** -53), 2 ** -60, 127.5 / 128,
               -127.5 / 128, 0.5 - 2 ** -54, (64 + 0.5 + 127.5 / 256) / 64]
    for x in list(r.uniform(-1, 1, 50)) + [v for v in special if abs(v) < 1]:
        sig = orc.split_value(float(x), s)
        assert list(sig.astype(int)) == _py_slices(float(x), s)
        assert abs(int(sig[0])) <= 65
        assert np.all(sig[1:].astype(int) >= -128) and np.all(sig[1:].astype(int) <= 127)
        # digits reproduce x to within half a unit of the last place 2^(2-8s)
        rec = sum(Fraction(int(v)) * Fraction(2) ** (2 - 8 * (p + 1)) for p, v in enumerate(sig))
        assert abs(Fraction(float(x)) - rec) <= Fraction(1, 2) * Fraction(2) ** (2 - 8 * s)


def test_split_value_carry_cases(orc):
    # 64 y = 1 + 127.75/256: the first digit is rint(1.499) = 1, the remainder
    # scaled by 256 is 127.75, whose rint is 128; the carry pass must turn the
    # digits [1, 128] into [2, -128].
    y = (1 + 127.75 / 256) / 64
    assert round((Fraction(y) * 64 - 1) * 256) == 128
    assert list(orc.split_value(y, 2).astype(int)) == [2, -128]
    # a chain of carries: 127.5 at every lower level rounds (half to even) to 128
    x = (0.5 + 127.5 / 256) / 64
    for s in (2, 3, 5):
        sig = orc.split_value(x, s).astype(int)
        assert list(sig) == _py_slices(x, s)
        assert all(-128 <= d <= 127 for d in sig[1:])


def test_split_exponents(orc):
    A = np.array([[0.75, -0.1], [0.0, 0.0], [-3.0, 2.5]])
    B = np.array([[1.0, 0.0], [-0.25, 0.0]])
    re, ce = orc.split_exponents(A, B)
    assert list(re) == [math.frexp(0.75)[1], 0, math.frexp(3.0)[1]]
    assert list(ce) == [math.frexp(1.0)[1], 0]


def test_split_chunk_respects_int32(orc):
    for s in range(1, 10):
        for t in range(2, s + 2):
            pairs = min(s, t - 1) - max(1, t - s) + 1
            c = orc.split_chunk_tiles(s, t)
            assert c % 2 == 0          # whole 128-byte stages
            assert pairs * c * 64 * 128 * 128 <= 2 ** 31 - 1
            assert pairs * (c + 2) * 64 * 128 * 128 > 2 ** 31 - 1


def _py_split_product(A, B, s):
    """Exact-arithmetic evaluation of the split definition (single K chunk)."""
    m, k = A.shape
    n = B.shape[1]
    re = [math.frexp(float(np.max(np.abs(A[i])) or 0.0))[1] if np.any(A[i]) else 0 for i in range(m)]
    ce = [math.frexp(float(np.max(np.abs(B[:, j]))))[1] if np.any(B[:, j]) else 0 for j in range(n)]
    SA = [[_py_slices(float(Fraction(A[i, t]) / Fraction(2) ** re[i]), s) for t in range(k)]
          for i in range(m)]
    SB = [[_py_slices(float(Fraction(B[t, j]) / Fraction(2) ** ce[j]), s) for t in range(k)]
          for j in range(n)]
    C = np.empty((m, n))
    for i in range(m):
        for j in range(n):
            v = None
            for t in range(2, s + 2):
                g = 0
                for p in range(max(1, t - s), min(s, t - 1) + 1):
                    q = t - p
                    g += sum(SA[i][kk][p - 1] * SB[j][kk][q - 1] for kk in range(k))
                x = fr.rnd(Fraction(g) * Fraction(2) ** (re[i] + ce[j] + 4 - 8 * t))
                v = x if v is None else fr.add(v, x)
            C[i, j] = v
    return C


@pytest.mark.parametrize("s", [1, 2, 5, 8])
def test_split_model_matches_exact_evaluation(orc, s):
    A = rng_matrix(4, 21, 81)
    B = rng_matrix(21, 3, 82)
    A[1] *= 2.0 ** 5
    B[:, 2] *= 2.0 ** -9
    assert np.array_equal(orc.split_gemm_int8(A, B, s), _py_split_product(A, B, s))


def test_split_model_exact_on_7bit_inputs(orc):
    # Entries with at most 7 significant bits relative to their row/column
    # maximum are represented exactly by the first two digits; with s >= 3 the
    # pair (2,2) is kept, so the product is exact.
    r = np.random.default_rng(5)
    A = r.integers(-127, 128, size=(31, 90)).astype(np.float64)
    B = r.integers(-127, 128, size=(90, 29)).astype(np.float64) * 2.0 ** -20
    exact = (A.astype(np.int64) @ (B * 2.0 ** 20).astype(np.int64)).astype(np.float64) * 2.0 ** -20
    for s in (3, 4, 8):
        assert np.array_equal(orc.split_gemm_int8(A, B, s), exact)


@pytest.mark.parametrize("s", [2, 4, 6, 8])
def test_split_model_error_bound(orc, s):
    # |C - AB| <= 2^(e_i+f_j) K [ (s+3) 2^(2-8s) + (s+2) u ]: rounding residuals
    # of both operands, the dropped pairs p+q > s+1, and the FP64 folds.
    m, n, k = 6, 5, 120
    A = rng_matrix(m, k, 91)
    B = rng_matrix(k, n, 92)
    A[2] *= 1e5
    C = orc.split_gemm_int8(A, B, s)
    re, ce = orc.split_exponents(A, B)
    u = 2.0 ** -53
    for i in range(m):
        for j in range(n):
            err = abs(float(Fraction(C[i, j]) - fr.exact_dot(A[i, :], B[:, j])))
            bound = 2.0 ** (int(re[i]) + int(ce[j])) * k * ((s + 3) * 2.0 ** (2 - 8 * s) + (s + 2) * u)
            assert err <= bound


def test_split_model_accumulate_and_nonfinite(orc):
    A = rng_matrix(5, 40, 101)
    B = rng_matrix(40, 6, 102)
    C0 = rng_matrix(5, 6, 103)
    # accumulate form: C_in is added before the first fold (exact here: with
    # 7-bit integer inputs every fold is exact)
    r = np.random.default_rng(8)
    Ai = r.integers(-127, 128, size=(5, 40)).astype(np.float64)
    Bi = r.integers(-127, 128, size=(40, 6)).astype(np.float64)
    Ci = r.integers(-1000, 1000, size=(5, 6)).astype(np.float64)
    acc = orc.split_gemm_int8(Ai, Bi, 6, C=Ci)
    assert np.array_equal(acc, Ci + Ai @ Bi)
    A[2, 7] = np.inf
    B[3, 4] = np.nan
    out = orc.split_gemm_int8(A, B, 6)
    assert np.all(np.isnan(out[2])) and np.all(np.isnan(out[:, 4]))
    ok = np.ones_like(out, dtype=bool)
    ok[2] = False
    ok[:, 4] = False
    assert np.all(np.isfinite(out[ok]))


# ---- cost-model MLP (double precision) ----------------------------------------------

def test_mlp_gradient_matches_finite_differences(orc):
    r = np.random.default_rng(0)
    D, n = 6, 40
    P = orc.mlp_param_count(D)
    assert P == D * 64 + 64 + 64 * 64 + 64 + 64 + 1
    th = r.normal(0, 0.3, P)
    X = r.normal(0, 1, (n, D))
    t = r.normal(0, 1, n)
    loss, g = orc.mlp_gradient(th, X, t)
    assert loss == pytest.approx(np.mean((orc.mlp_forward(th, X) - t) ** 2), rel=1e-12)
    idx = r.choice(P, 60, replace=False)
    idx = np.concatenate([idx, [P - 1, D * 64, D * 64 + 64 + 5, P - 2]])  # b3, b1, W2, W3 entries
    h = 1e-6
    for i in idx:
        tp = th.copy(); tp[i] += h
        tm = th.copy(); tm[i] -= h
        fd = (orc.mlp_gradient(tp, X, t)[0] - orc.mlp_gradient(tm, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_mlp_forward_closed_form(orc):
    # With W2 = I-like positive pass-through, the network reduces to a linear
    # model on positive inputs: y = b3 + sum_j W3_j (b2_j + relu(b1_j + x W1_j)).
    D = 2
    P = orc.mlp_param_count(D)
    th = np.zeros(P)
    o_b1, o_w2, o_b2, o_w3, o_b3 = D * 64, D * 64 + 64, D * 64 + 64 + 4096, D * 64 + 64 + 4096 + 64, P - 1
    th[0] = 2.0          # W1[0][0]
    th[64 + 0] = 3.0     # W1[1][0]
    th[o_w2 + 0] = 1.0   # W2[0][0]
    th[o_w3 + 0] = 0.5
    th[o_b3] = 1.0
    X = np.array([[1.0, 1.0], [2.0, -1.0], [-5.0, 0.0]])
    assert np.allclose(orc.mlp_forward(th, X), [1.0 + 0.5 * 5.0, 1.0 + 0.5 * 1.0, 1.0])
