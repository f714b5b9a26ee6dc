/*
 * oracle.c -- CPU oracle for the local FP64 block product.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the package's smoke
 * check and bench.py's CPU-baseline leg load this library, and only as the
 * checker (or, for the baseline, as the timed CPU restatement of the
 * reference's product).  The product path (libmgx) never calls it.
 *
 * Contents
 *  - orc_splitmix64_at / orc_fill*: the synthetic-input definition shared
 *    with libmgx's fill kernels (splitmix64 counter stream).
 *  - orc_gemm_kernel_order: restatement of the reference GEMM arithmetic,
 *    sharedABMultiply (1D-row-block/mulmat_kernel.cu:39-80): every C element
 *    is one sequential chain of fused multiply-adds over k = 0..K-1 starting
 *    from +0.0 (nvcc contracts `sum += a*b` into DFMA; out-of-range tile
 *    lanes add 0*0, which leaves the chain unchanged), then stored.
 *  - orc_gemm_accumulate_order: restatement of sharedABMultiplyAccumulate
 *    (1D-row-block/mulmat_kernel.cu:100-142, 2D-block-matrix/mulmat_kernel.cu:
 *    94-136): a chunk partial by the same chain over kk < k_chunk with
 *    k_offset + kk < k_total, then C += partial once.
 *  - orc_gemm_dd: double-double dot products (Ogita-Rump-Oishi Dot2 with
 *    Dekker/Veltkamp TwoProduct realised by fma and Knuth TwoSum), giving an
 *    accuracy reference far below FP64 rounding.
 *  - orc_validate_constant: the reference validation rule
 *    (1D-row-block/mulmat_kernel.cu:180-187) and the 2D host rule
 *    (2D-block-matrix/mmb.cu:381-385).
 *  - orc_split_*: a bit-exact model of libmgx's INT8 split-precision
 *    (Ozaki scheme I) product: exponents, slices, exact group sums with the
 *    same INT32 flush chunks, and the same FP64 fold sequence into C.
 *  - orc_crt_*: a bit-exact model of libmgx's INT8 modular (Ozaki scheme II)
 *    product: the same exponents, the integer scaling by 2^beta, the exact
 *    integer product in 128-bit arithmetic and one correctly rounded
 *    conversion per output.
 *
 * Parity pins.  tests/test_oracle.py: exact rational brute force on tiny
 * inputs for the fma chains, the accumulate form, the double-double dot and
 * the split reconstruction; the A=1/B=2 known answer; integer-valued inputs
 * (exact in every order); transposition identity; the splitmix64 published
 * first outputs for seed 0; closed-form validation cases; the MLP against
 * finite differences and a closed form.  tests/test_oracle_crt.py: the
 * orc_crt_* model against Python arbitrary-precision integers, the scaling
 * rule's inequality in exact rationals and the classical exponent on dense
 * lines.  tests/test_oracle_generators.py: the stress-input generators
 * (exponent range, exact power-of-two scaling) in closed form.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

ORC_API int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- synthetic inputs ---------------------------------------------------- */

ORC_API uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

ORC_API void orc_fill(double *X, int64_t rows, int64_t cols, int64_t ld, int kind, uint64_t seed,
                      double value) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r) {
        for (int64_t c = 0; c < cols; ++c) {
            const uint64_t idx = (uint64_t)(r * cols + c);
            double v;
            if (kind == 0) {
                v = value;
            } else if (kind == 1) {
                v = 2.0 * (double)(orc_splitmix64_at(seed, idx + 1) >> 11) * 0x1.0p-53 - 1.0;
            } else {
                const uint64_t x = orc_splitmix64_at(seed, 2 * idx + 1);
                const uint64_t y = orc_splitmix64_at(seed, 2 * idx + 2);
                const double mant = 1.0 + (double)(x >> 12) * 0x1.0p-52;
                const int e = (int)((y >> 11) % 61ull) - 30;
                v = ldexp((y & 1ull) ? -mant : mant, e);
            }
            X[r * ld + c] = v;
        }
    }
}

ORC_API void orc_fill_exponents(int32_t *e, int64_t count, uint64_t seed, int lo, int hi) {
    const uint64_t span = (uint64_t)(hi - lo + 1);
    for (int64_t i = 0; i < count; ++i)
        e[i] = lo + (int)((orc_splitmix64_at(seed, (uint64_t)i + 1) >> 11) % span);
}

ORC_API void orc_scale_rows_cols(double *X, int64_t rows, int64_t cols, int64_t ld,
                                 const int32_t *re, const int32_t *ce) {
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < cols; ++c)
            X[r * ld + c] = ldexp(X[r * ld + c], (re ? re[r] : 0) + (ce ? ce[c] : 0));
}

/* ---- reference-order GEMM ------------------------------------------------ */

/* Per-element chain: acc = fma(A[i][kk], B[kk][j], acc) for kk = 0..kd-1,
 * starting from +0.0.  Rows are blocked (RB) and columns tiled (JB) so B is
 * streamed once per row block; the per-element order is unchanged. */
#define ORC_RB 16
#define ORC_JB 512
static void chain_block(const double *A, const double *B, double *P, int64_t m, int64_t n,
                        int64_t kd, int64_t lda, int64_t ldb, int64_t ldp) {
    const int64_t nrb = (m + ORC_RB - 1) / ORC_RB;
    const int64_t njb = (n + ORC_JB - 1) / ORC_JB;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t rb = 0; rb < nrb; ++rb) {
        for (int64_t jb = 0; jb < njb; ++jb) {
            const int64_t i0 = rb * ORC_RB, i1 = i0 + ORC_RB < m ? i0 + ORC_RB : m;
            const int64_t j0 = jb * ORC_JB, j1 = j0 + ORC_JB < n ? j0 + ORC_JB : n;
            const int64_t w = j1 - j0;
            double acc[ORC_RB][ORC_JB];
            for (int64_t i = i0; i < i1; ++i)
                for (int64_t j = 0; j < w; ++j) acc[i - i0][j] = 0.0;
            for (int64_t kk = 0; kk < kd; ++kk) {
                const double *brow = B + kk * ldb + j0;
                for (int64_t i = i0; i < i1; ++i) {
                    const double a = A[i * lda + kk];
                    double *ar = acc[i - i0];
                    for (int64_t j = 0; j < w; ++j) ar[j] = fma(a, brow[j], ar[j]);
                }
            }
            for (int64_t i = i0; i < i1; ++i)
                for (int64_t j = 0; j < w; ++j) P[i * ldp + j0 + j] = acc[i - i0][j];
        }
    }
}

ORC_API void orc_gemm_kernel_order(const double *A, const double *B, double *C, int64_t m,
                                   int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                                   int nthreads) {
    set_threads(nthreads);
    chain_block(A, B, C, m, n, k, lda, ldb, ldc);
}

ORC_API int orc_gemm_accumulate_order(const double *A, const double *Bc, double *C, int64_t m,
                                      int64_t n, int64_t k_total, int64_t k_offset,
                                      int64_t k_chunk, int64_t lda, int64_t ldb, int64_t ldc,
                                      int nthreads) {
    set_threads(nthreads);
    int64_t kd = k_total - k_offset;
    if (kd > k_chunk) kd = k_chunk;
    if (kd < 0) kd = 0;
    double *P = (double *)malloc(sizeof(double) * (size_t)(m * n > 0 ? m * n : 1));
    if (!P) return -1;
    chain_block(A + k_offset, Bc, P, m, n, kd, lda, ldb, n);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) C[i * ldc + j] = C[i * ldc + j] + P[i * n + j];
    free(P);
    return 0;
}

/* ---- double-double reference --------------------------------------------- */

/* hi + lo ~ sum_k A[i][k] B[k][j] with Dot2: TwoProduct(a,b) = (a*b, fma(a,b,-a*b)),
 * TwoSum(x,y) = Knuth's branch-free six-operation sum.  Only the rows listed
 * in `rows` (nrows entries) are computed; output row r of Chi/Clo holds
 * source row rows[r]. */
ORC_API void orc_gemm_dd_rows(const double *A, const double *B, double *Chi, double *Clo,
                              const int64_t *rows, int64_t nrows, int64_t n, int64_t k,
                              int64_t lda, int64_t ldb, int64_t ldc, int nthreads) {
    set_threads(nthreads);
    const int64_t njb = (n + ORC_JB - 1) / ORC_JB;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t r = 0; r < nrows; ++r) {
        for (int64_t jb = 0; jb < njb; ++jb) {
            const int64_t i = rows[r];
            const int64_t j0 = jb * ORC_JB, j1 = j0 + ORC_JB < n ? j0 + ORC_JB : n;
            const int64_t w = j1 - j0;
            double p[ORC_JB], s[ORC_JB];
            for (int64_t j = 0; j < w; ++j) p[j] = s[j] = 0.0;
            for (int64_t kk = 0; kk < k; ++kk) {
                const double a = A[i * lda + kk];
                const double *brow = B + kk * ldb + j0;
                for (int64_t j = 0; j < w; ++j) {
                    const double h = a * brow[j];
                    const double e = fma(a, brow[j], -h);
                    const double x = p[j] + h;
                    const double z = x - p[j];
                    const double q = (p[j] - (x - z)) + (h - z);
                    p[j] = x;
                    s[j] = s[j] + (q + e);
                }
            }
            for (int64_t j = 0; j < w; ++j) {
                /* renormalise: hi = fl(p + s), lo = error of that sum */
                const double x = p[j] + s[j];
                const double z = x - p[j];
                const double lo = (p[j] - (x - z)) + (s[j] - z);
                Chi[r * ldc + j0 + j] = x;
                Clo[r * ldc + j0 + j] = lo;
            }
        }
    }
}

/* |A|.|B| for the rows listed (componentwise-error denominators). */
ORC_API void orc_gemm_abs_rows(const double *A, const double *B, double *D, const int64_t *rows,
                               int64_t nrows, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                               int64_t ldd, int nthreads) {
    set_threads(nthreads);
    const int64_t njb = (n + ORC_JB - 1) / ORC_JB;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t r = 0; r < nrows; ++r) {
        for (int64_t jb = 0; jb < njb; ++jb) {
            const int64_t i = rows[r];
            const int64_t j0 = jb * ORC_JB, j1 = j0 + ORC_JB < n ? j0 + ORC_JB : n;
            const int64_t w = j1 - j0;
            double acc[ORC_JB];
            for (int64_t j = 0; j < w; ++j) acc[j] = 0.0;
            for (int64_t kk = 0; kk < k; ++kk) {
                const double a = fabs(A[i * lda + kk]);
                const double *brow = B + kk * ldb + j0;
                for (int64_t j = 0; j < w; ++j) acc[j] = fma(a, fabs(brow[j]), acc[j]);
            }
            for (int64_t j = 0; j < w; ++j) D[r * ldd + j0 + j] = acc[j];
        }
    }
}

/* ---- validation rules ------------------------------------------------------ */

/* 1D rule: invalid iff |C-e| > abs_tol && |C-e|/|e| > rel_tol (rel = 0 when
 * e == 0).  Max errors ignore NaN (fmax). Returns the invalid count. */
ORC_API uint64_t orc_validate_constant(const double *C, int64_t rows, int64_t cols, int64_t ld,
                                       double expected, double abs_tol, double rel_tol,
                                       double *max_abs, double *max_rel) {
    uint64_t bad = 0;
    double ma = 0.0, mr = 0.0;
    const double ae = fabs(expected);
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < cols; ++c) {
            const double abs_err = fabs(C[r * ld + c] - expected);
            const double rel_err = ae > 0.0 ? abs_err / ae : 0.0;
            if ((abs_err > abs_tol) && (rel_err > rel_tol)) ++bad;
            ma = fmax(ma, abs_err);
            mr = fmax(mr, rel_err);
        }
    if (max_abs) *max_abs = ma;
    if (max_rel) *max_rel = mr;
    return bad;
}

/* 2D host rule: valid iff every element is finite and |C-e| <= 1e-9 + 1e-12|e|.
 * Returns the first failing linear index, or -1 when all pass. */
ORC_API int64_t orc_validate_2d(const double *C, int64_t count, double expected) {
    const double tol = 1e-9 + 1e-12 * fabs(expected);
    for (int64_t i = 0; i < count; ++i)
        if (!isfinite(C[i]) || fabs(C[i] - expected) > tol) return i;
    return -1;
}

/* ---- split-precision model (INT8 slices, rounded signed digits) ------------ */

#define ORC_EXP_BAD 0x40000000
#define ORC_KT 64
#define ORC_KALIGN 256

static int exponent_of_max(double mx, int bad) {
    if (bad) return ORC_EXP_BAD;
    if (mx == 0.0) return 0;
    int e;
    frexp(mx, &e);
    return e;
}

/* Row exponents of A (m x k) and column exponents of B (k x n). */
ORC_API void orc_split_exponents(const double *A, const double *B, int64_t m, int64_t n,
                                 int64_t k, int64_t lda, int64_t ldb, int32_t *rexp,
                                 int32_t *cexp) {
    for (int64_t i = 0; i < m; ++i) {
        double mx = 0.0;
        int bad = 0;
        for (int64_t kk = 0; kk < k; ++kk) {
            const double a = A[i * lda + kk];
            if (!isfinite(a)) bad = 1;
            else mx = fmax(mx, fabs(a));
        }
        rexp[i] = exponent_of_max(mx, bad);
    }
    for (int64_t j = 0; j < n; ++j) {
        double mx = 0.0;
        int bad = 0;
        for (int64_t kk = 0; kk < k; ++kk) {
            const double b = B[kk * ldb + j];
            if (!isfinite(b)) bad = 1;
            else mx = fmax(mx, fabs(b));
        }
        cexp[j] = exponent_of_max(mx, bad);
    }
}

/* Digits of one scaled value x in (-1,1): sig[p-1], p = 1..s, place value
 * 2^(2-8p).  sigma_1 = rint(x 2^6); sigma_p = rint(256 r) of the previous
 * remainder; then carries from the least significant digit turn 128 into
 * -128 (+1 on the digit above). */
ORC_API void orc_split_value(double x, int s, int8_t *sig) {
    int dig[16];
    double y = x * 64.0;
    for (int p = 0; p < s; ++p) {
        const double d = rint(y);
        dig[p] = (int)d;
        y = (y - d) * 256.0;
    }
    for (int p = s - 1; p > 0; --p)
        if (dig[p] >= 128) {
            dig[p] -= 256;
            dig[p - 1] += 1;
        }
    for (int p = 0; p < s; ++p) sig[p] = (int8_t)dig[p];
}

/* Group-t chunk length in 64-deep tiles, the INT32 bound of the product kernel. */
ORC_API int orc_split_chunk_tiles(int s, int t) {
    const int pmin = t - s > 1 ? t - s : 1;
    const int pmax = t - 1 < s ? t - 1 : s;
    const long long np = pmax - pmin + 1;
    long long c = 2147483647ll / (np * 128ll * 128ll * ORC_KT);
    /* whole 128-byte LDS stages of the product kernel */
    c = (c / 2 < 1 ? 1 : c / 2) * 2;
    return (int)c;
}

/* Full model: C = A.B (or C += A.B) exactly as the INT8 split kernel does. */
ORC_API int orc_split_gemm_int8(int s, const double *A, const double *B, double *C, int64_t m,
                                int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                                int accumulate, int nthreads) {
    set_threads(nthreads);
    if (s < 1 || s > 9) return -1;
    if (m == 0 || n == 0) return 0;
    if (k == 0) {
        if (!accumulate)
            for (int64_t i = 0; i < m; ++i)
                for (int64_t j = 0; j < n; ++j) C[i * ldc + j] = 0.0;
        return 0;
    }
    int32_t *rexp = (int32_t *)malloc(sizeof(int32_t) * m);
    int32_t *cexp = (int32_t *)malloc(sizeof(int32_t) * n);
    int8_t *As = (int8_t *)malloc((size_t)s * m * k);
    int8_t *Bs = (int8_t *)malloc((size_t)s * n * k); /* transposed: [q][j][kk] */
    if (!rexp || !cexp || !As || !Bs) {
        free(rexp); free(cexp); free(As); free(Bs);
        return -2;
    }
    orc_split_exponents(A, B, m, n, k, lda, ldb, rexp, cexp);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t kk = 0; kk < k; ++kk) {
            int8_t sig[16];
            const double x = rexp[i] == ORC_EXP_BAD ? 0.0 : ldexp(A[i * lda + kk], -rexp[i]);
            orc_split_value(x, s, sig);
            for (int p = 0; p < s; ++p) As[((size_t)p * m + i) * k + kk] = sig[p];
        }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t kk = 0; kk < k; ++kk) {
            int8_t sig[16];
            const double x = cexp[j] == ORC_EXP_BAD ? 0.0 : ldexp(B[kk * ldb + j], -cexp[j]);
            orc_split_value(x, s, sig);
            for (int p = 0; p < s; ++p) Bs[((size_t)p * n + j) * k + kk] = sig[p];
        }
    const int64_t ktiles = ((k + ORC_KALIGN - 1) / ORC_KALIGN * ORC_KALIGN) / ORC_KT;
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 1) collapse(2) reduction(| : overflow)
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            /* Fold sequence of the product kernel: x = ldexp(G, e_i+f_j+4-8t) per
             * (group, chunk); C = x (or C_in + x) first, then C = C + x. */
            double v = accumulate ? C[i * ldc + j] : 0.0;
            int first = 1;
            const int bad = rexp[i] == ORC_EXP_BAD || cexp[j] == ORC_EXP_BAD;
            for (int t = 2; t <= s + 1; ++t) {
                const int pmin = t - s > 1 ? t - s : 1;
                const int pmax = t - 1 < s ? t - 1 : s;
                const int64_t chunk = orc_split_chunk_tiles(s, t);
                for (int64_t c0 = 0; c0 < ktiles; c0 += chunk) {
                    int64_t k0 = c0 * ORC_KT;
                    int64_t k1 = (c0 + chunk) * ORC_KT;
                    if (k1 > k) k1 = k;
                    long long g = 0;
                    for (int p = pmin; p <= pmax; ++p) {
                        const int8_t *ar = As + ((size_t)(p - 1) * m + i) * k;
                        const int8_t *br = Bs + ((size_t)(t - p - 1) * n + j) * k;
                        for (int64_t kk = k0; kk < k1; ++kk) g += (long long)ar[kk] * br[kk];
                    }
                    if (g > 2147483647ll || g < -2147483648ll) overflow = 1;
                    const double x = bad ? NAN : ldexp((double)g, rexp[i] + cexp[j] + 4 - 8 * t);
                    if (first)
                        v = accumulate ? v + x : x;
                    else
                        v = v + x;
                    first = 0;
                }
            }
            C[i * ldc + j] = v;
        }
    }
    free(rexp); free(cexp); free(As); free(Bs);
    return overflow ? -3 : 0;
}

/* ---- INT8 modular form (Ozaki scheme II) ---------------------------------- */
/* Published algorithm: K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A
 * GEMM-oriented emulation of floating-point matrix multiplication using an
 * integer modular technique" (2025).  Row i is scaled to integers
 * A'_ik = rint(A_ik 2^(beta_i-e_i)) (|A'| <= 2^beta_i), column j of B likewise
 * with f_j and delta_j; the integer product C' = A'B' is formed exactly
 * (libmgx forms it from INT8 products modulo n pairwise coprime moduli and
 * Chinese remaindering; this model forms it directly in 128-bit integers,
 * which is the same integer whenever |C'| < P/2, P = m_0 ... m_{n-1}, which
 * the per-line exponents guarantee), and
 *   C_ij = ldexp((double)C'_ij, e_i - beta_i + f_j - delta_j)   (one rounding)
 * (or C_in + that for the accumulate form).
 *
 * Per-line exponent (this build's own rule; Cauchy-Schwarz on an exponent
 * histogram): for a line x_1..x_k (a row of A or a column of B) with e(x) the
 * frexp exponent and groups of 16 consecutive k,
 *   N = sum_groups sum_{x != 0 in group} 4^(12 - min(g - e(x), 12)),
 *   g = the group's largest e(x);
 * R = floor(sqrt(P/2)) - ceil(ceil(sqrt(k))/2) - 1; beta is the largest
 * integer in [0, 53] with 4^beta N <= R^2 4^12.  Then sum x^2 < 4^e N/4^12,
 * ||X'||_2 < R + sqrt(k)/2 and |C'_ij| <= ||A'_i|| ||B'_j|| < P/2. */
#define ORC_MAX_MODULI 16
static const uint32_t orc_moduli[ORC_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233,
                                                     229, 227, 223, 217, 211, 199, 197, 193};

ORC_API int orc_crt_modulus(int l) { return (l >= 0 && l < ORC_MAX_MODULI) ? (int)orc_moduli[l] : -1; }

static unsigned __int128 orc_crt_product(int nmod) {
    unsigned __int128 P = 1;
    for (int l = 0; l < nmod; ++l) P *= orc_moduli[l];
    return P;
}

/* log2 of P_n (for reporting). */
ORC_API double orc_crt_log2_product(int nmod) {
    double s = 0.0;
    for (int l = 0; l < nmod && l < ORC_MAX_MODULI; ++l) s += log2((double)orc_moduli[l]);
    return s;
}

#define ORC_HIST_W 12
#define ORC_GROUP_K 16

static unsigned __int128 orc_isqrt(unsigned __int128 x) {
    unsigned __int128 r = (unsigned __int128)sqrtl((long double)x);
    while (r > 0 && r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r;
}

/* R^2 of the per-line rule, 0 when R < 1. */
static unsigned __int128 orc_crt_r2(int nmod, int64_t k) {
    const unsigned __int128 P = orc_crt_product(nmod);
    const unsigned __int128 s = orc_isqrt(P / 2);
    unsigned __int128 h = orc_isqrt((unsigned __int128)k);
    if (h * h < (unsigned __int128)k) h += 1;
    const unsigned __int128 cut = (h + 1) / 2 + 1;
    if (s <= cut) return 0;
    return (s - cut) * (s - cut);
}

/* Largest beta in [0, 53] with 4^beta N <= R2 4^12, or -1. */
static int orc_crt_beta_of(unsigned __int128 R2, uint64_t N) {
    for (int b = 53; b >= 0; --b) {
        /* 4^b N <= R2 4^12  <=>  N 4^(b-12) <= R2 (b >= 12), N <= R2 4^(12-b) */
        if (b >= ORC_HIST_W) {
            /* N 2^sh <= R2  <=>  N <= floor(R2 / 2^sh) */
            if ((unsigned __int128)N <= (R2 >> (2 * (b - ORC_HIST_W)))) return b;
        } else {
            /* N <= R2 2^sh  <=>  ceil(N / 2^sh) <= R2 */
            const int sh = 2 * (ORC_HIST_W - b);
            if ((((unsigned __int128)N + ((unsigned __int128)1 << sh) - 1) >> sh) <= R2) return b;
        }
    }
    return -1;
}

/* Exponent of the densest line (every element in the top binade,
 * N = k 4^12): the exponent every line gets at least. */
ORC_API int orc_crt_beta(int nmod, int64_t k) {
    if (nmod < 2 || nmod > ORC_MAX_MODULI || k < 1 || k > ((int64_t)1 << 36)) return -1;
    const unsigned __int128 R2 = orc_crt_r2(nmod, k);
    if (R2 == 0) return -1;
    return orc_crt_beta_of(R2, (uint64_t)k << (2 * ORC_HIST_W));
}

/* N of one line: x(i) = X[i * kstride], i < k. */
static uint64_t orc_crt_line_sum(const double *X, int64_t k, int64_t kstride) {
    uint64_t N = 0;
    for (int64_t g0 = 0; g0 < k; g0 += ORC_GROUP_K) {
        const int64_t g1 = g0 + ORC_GROUP_K < k ? g0 + ORC_GROUP_K : k;
        int top = INT32_MIN;
        for (int64_t i = g0; i < g1; ++i) {
            const double x = X[i * kstride];
            if (x != 0.0 && isfinite(x)) {
                int e;
                frexp(x, &e);
                if (e > top) top = e;
            }
        }
        for (int64_t i = g0; i < g1; ++i) {
            const double x = X[i * kstride];
            if (x != 0.0 && isfinite(x)) {
                int e;
                frexp(x, &e);
                const int d = top - e < ORC_HIST_W ? top - e : ORC_HIST_W;
                N += (uint64_t)1 << (2 * (ORC_HIST_W - d));
            }
        }
    }
    return N;
}

/* Per-line exponents of `lines` lines: line l is X[l * lstride + i * kstride],
 * i < k.  Returns 0, or -1 when the moduli cannot hold depth k. */
ORC_API int orc_crt_line_betas(int nmod, const double *X, int64_t lines, int64_t k,
                               int64_t lstride, int64_t kstride, int32_t *beta) {
    if (orc_crt_beta(nmod, k) < 0) return -1;
    const unsigned __int128 R2 = orc_crt_r2(nmod, k);
    for (int64_t l = 0; l < lines; ++l)
        beta[l] = orc_crt_beta_of(R2, orc_crt_line_sum(X + l * lstride, k, kstride));
    return 0;
}

ORC_API int orc_crt_gemm(int nmod, const double *A, const double *B, double *C, int64_t m,
                         int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                         int accumulate, int nthreads) {
    set_threads(nthreads);
    if (m == 0 || n == 0) return 0;
    if (k == 0) {
        if (!accumulate)
            for (int64_t i = 0; i < m; ++i)
                for (int64_t j = 0; j < n; ++j) C[i * ldc + j] = 0.0;
        return 0;
    }
    if (orc_crt_beta(nmod, k) < 0) return -1;
    int32_t *rexp = (int32_t *)malloc(sizeof(int32_t) * m);
    int32_t *cexp = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *rbeta = (int32_t *)malloc(sizeof(int32_t) * m);
    int32_t *cbeta = (int32_t *)malloc(sizeof(int32_t) * n);
    int64_t *Ai = (int64_t *)malloc(sizeof(int64_t) * m * k);
    int64_t *Bi = (int64_t *)malloc(sizeof(int64_t) * n * k); /* transposed: [j][kk] */
    if (!rexp || !cexp || !rbeta || !cbeta || !Ai || !Bi) {
        free(rexp); free(cexp); free(rbeta); free(cbeta); free(Ai); free(Bi);
        return -2;
    }
    orc_split_exponents(A, B, m, n, k, lda, ldb, rexp, cexp);
    orc_crt_line_betas(nmod, A, m, k, lda, 1, rbeta);
    orc_crt_line_betas(nmod, B, n, k, 1, ldb, cbeta);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t kk = 0; kk < k; ++kk)
            Ai[i * k + kk] = rexp[i] == ORC_EXP_BAD
                                 ? 0
                                 : (int64_t)rint(ldexp(A[i * lda + kk], rbeta[i] - rexp[i]));
    for (int64_t j = 0; j < n; ++j)
        for (int64_t kk = 0; kk < k; ++kk)
            Bi[j * k + kk] = cexp[j] == ORC_EXP_BAD
                                 ? 0
                                 : (int64_t)rint(ldexp(B[kk * ldb + j], cbeta[j] - cexp[j]));
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            const int64_t *ar = Ai + i * k;
            const int64_t *br = Bi + j * k;
            __int128 acc = 0;
            for (int64_t kk = 0; kk < k; ++kk) acc += (__int128)ar[kk] * br[kk];
            const int bad = rexp[i] == ORC_EXP_BAD || cexp[j] == ORC_EXP_BAD;
            const double x =
                bad ? NAN : ldexp((double)acc, rexp[i] - rbeta[i] + cexp[j] - cbeta[j]);
            C[i * ldc + j] = accumulate ? C[i * ldc + j] + x : x;
        }
    }
    free(rexp); free(cexp); free(rbeta); free(cbeta); free(Ai); free(Bi);
    return 0;
}

/* ---- cost-model MLP (double precision) ------------------------------------ */
/* Same network and parameter layout as libmgx's FP32 MLP: W1 [D][64], b1,
 * W2 [64][64], b2, W3 [64], b3; loss = mean (y - t)^2.  Evaluated in double
 * so the GPU's FP32 gradients can be checked against it; pinned itself by
 * central finite differences. */
#define ORC_H 64
static void mlp_offsets(int D, int64_t *w1, int64_t *b1, int64_t *w2, int64_t *b2, int64_t *w3,
                        int64_t *b3) {
    *w1 = 0;
    *b1 = (int64_t)D * ORC_H;
    *w2 = *b1 + ORC_H;
    *b2 = *w2 + ORC_H * ORC_H;
    *w3 = *b2 + ORC_H;
    *b3 = *w3 + ORC_H;
}

ORC_API int64_t orc_mlp_param_count(int D) { return (int64_t)D * ORC_H + ORC_H * ORC_H + 3 * ORC_H + 1; }

ORC_API void orc_mlp_forward(const double *th, int D, const double *X, int64_t n, double *y) {
    int64_t w1, b1, w2, b2, w3, b3;
    mlp_offsets(D, &w1, &b1, &w2, &b2, &w3, &b3);
    for (int64_t s = 0; s < n; ++s) {
        double h1[ORC_H], out = th[b3];
        for (int j = 0; j < ORC_H; ++j) {
            double a = th[b1 + j];
            for (int d = 0; d < D; ++d) a += X[s * D + d] * th[w1 + d * ORC_H + j];
            h1[j] = a > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < ORC_H; ++j) {
            double a = th[b2 + j];
            for (int i = 0; i < ORC_H; ++i) a += h1[i] * th[w2 + i * ORC_H + j];
            out += (a > 0.0 ? a : 0.0) * th[w3 + j];
        }
        y[s] = out;
    }
}

/* Loss (returned) and gradient (g, param_count entries) of the mean squared error. */
ORC_API double orc_mlp_gradient(const double *th, int D, const double *X, const double *t,
                                int64_t n, double *g) {
    int64_t w1, b1, w2, b2, w3, b3;
    mlp_offsets(D, &w1, &b1, &w2, &b2, &w3, &b3);
    const int64_t P = orc_mlp_param_count(D);
    for (int64_t p = 0; p < P; ++p) g[p] = 0.0;
    double loss = 0.0;
    for (int64_t s = 0; s < n; ++s) {
        double a1[ORC_H], h1[ORC_H], a2[ORC_H], h2[ORC_H], y = th[b3];
        for (int j = 0; j < ORC_H; ++j) {
            double a = th[b1 + j];
            for (int d = 0; d < D; ++d) a += X[s * D + d] * th[w1 + d * ORC_H + j];
            a1[j] = a;
            h1[j] = a > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < ORC_H; ++j) {
            double a = th[b2 + j];
            for (int i = 0; i < ORC_H; ++i) a += h1[i] * th[w2 + i * ORC_H + j];
            a2[j] = a;
            h2[j] = a > 0.0 ? a : 0.0;
            y += h2[j] * th[w3 + j];
        }
        const double diff = y - t[s];
        loss += diff * diff / (double)n;
        const double dy = 2.0 * diff / (double)n;
        g[b3] += dy;
        double dh2[ORC_H];
        for (int j = 0; j < ORC_H; ++j) {
            g[w3 + j] += h2[j] * dy;
            dh2[j] = a2[j] > 0.0 ? th[w3 + j] * dy : 0.0;
            g[b2 + j] += dh2[j];
            for (int i = 0; i < ORC_H; ++i) g[w2 + i * ORC_H + j] += h1[i] * dh2[j];
        }
        for (int i = 0; i < ORC_H; ++i) {
            double back = 0.0;
            for (int k = 0; k < ORC_H; ++k) back += th[w2 + i * ORC_H + k] * dh2[k];
            const double dh1 = a1[i] > 0.0 ? back : 0.0;
            g[b1 + i] += dh1;
            for (int d = 0; d < D; ++d) g[w1 + d * ORC_H + i] += X[s * D + d] * dh1;
        }
    }
    return loss;
}
