// Modular arithmetic of the INT8 modular split form (Ozaki scheme II):
// the moduli, the Chinese-remainder tables, residue extraction and the exact
// reconstruction of each output.  Internal to libmgx.
//
// Published algorithm: K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique" (2025).  Row i of A is scaled to integers
// A' = rint(A 2^(beta_i-e_i)), |A'| <= 2^beta_i, column j of B likewise with
// its own delta_j; for each modulus m_l the INT8 matrix cores form
// (A' mod m_l)(B' mod m_l) exactly in INT32, reduced mod m_l; Chinese
// remaindering over the n moduli recovers the integer C' = A'B' exactly
// because every |C'_ij| < P/2, P = m_0 ... m_{n-1} (the per-line scaling rule
// below guarantees it).  The output is ldexp(RN(C'), e_i - beta_i + f_j -
// delta_j): one rounding.
#pragma once

#include <stdint.h>

namespace mgx {
namespace modular {

constexpr int MIN_MODULI = 2;
constexpr int MAX_MODULI = 16;

// Pairwise coprime, at most 256 (so residues fit int8 and products of two
// residues fit the INT8 MFMA), largest first; P_16 < 2^127.
constexpr uint32_t MODULI[MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233,
                                         229, 227, 223, 217, 211, 199, 197, 193};

constexpr uint32_t inverse_mod(uint32_t a, uint32_t m) {
    // extended Euclid on (a mod m, m); a and m coprime
    int64_t r0 = m, r1 = a % m, t0 = 0, t1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1;
        const int64_t r2 = r0 - q * r1;
        r0 = r1;
        r1 = r2;
        const int64_t t2 = t0 - q * t1;
        t0 = t1;
        t1 = t2;
    }
    return (uint32_t)(((t0 % (int64_t)m) + m) % m);
}

// ---- per-line scaling exponents ------------------------------------------
//
// By Cauchy-Schwarz |C'_ij| <= ||A'_i||_2 ||B'_j||_2, so each line (a row of
// A, a column of B) gets its own scale 2^(beta - E), E = the frexp exponent of
// the line's largest |x|, with beta as large as keeps ||X'||_2 below a radius
// R where (R + sqrt(k)/2)^2 < P/2 (sqrt(k)/2 bounds the rounding of the
// scaled line to integers).  The 2-norm is bounded from exponents alone, so
// the bound is an exact integer that no summation order changes: with e(x)
// the frexp exponent of x (|x| < 2^e(x)),
//   N = sum over groups of GROUP_K consecutive k, over the nonzero x of the
//       group, of 4^(W - min(g - e(x), W)),   g = the group's largest e(x),
// and sum |x|^2 < 4^E N / 4^W, because every term relative to the group's
// maximum is at least the term relative to the line's (E >= g), and an
// element more than W binades below its group's maximum still counts 4^0.
// Then ||X 2^(beta - E)||_2 < 2^beta sqrt(N / 4^W) <= R whenever
// 4^beta N <= R^2 4^W, and beta is the largest such value in [0, 53].  A line
// whose elements all sit in its top binade has N = k 4^W and gets the
// classical worst-case exponent (largest beta with 2 k 4^beta < P, up to the
// rounding margin); lines with a spread of magnitudes get more bits.
constexpr int HIST_W = 12;       // binades resolved below a group maximum
constexpr int GROUP_K = 16;      // k values per group
constexpr int BETA_MAX = 53;

// Largest N that allows beta = b, for b = 0..53 (non-increasing in b).
struct BetaThresholds {
    unsigned long long t[BETA_MAX + 1];
};

inline unsigned __int128 isqrt_u128(unsigned __int128 x) {
    // floor(sqrt(x)) by bisection on the result's bits
    unsigned __int128 r = 0;
    for (int b = 63; b >= 0; --b) {
        const unsigned __int128 c = r | ((unsigned __int128)1 << b);
        if (c * c <= x) r = c;
    }
    return r;
}

// R^2 for n moduli at depth k: R = floor(sqrt(P/2)) - ceil(ceil(sqrt(k)) / 2) - 1,
// so R + sqrt(k)/2 < sqrt(P/2); 0 when R < 1 or the arguments are out of range.
inline unsigned __int128 radius2(int n, int64_t k) {
    if (n < MIN_MODULI || n > MAX_MODULI || k < 1) return 0;
    unsigned __int128 P = 1;
    for (int l = 0; l < n; ++l) P *= MODULI[l];
    const unsigned __int128 s = isqrt_u128(P / 2);
    unsigned __int128 h = isqrt_u128((unsigned __int128)k);
    if (h * h < (unsigned __int128)k) ++h;
    const unsigned __int128 cut = (h + 1) / 2 + 1;
    if (s <= cut) return 0;
    const unsigned __int128 R = s - cut;
    return R * R;
}

// Thresholds of the per-line rule: beta = b is allowed iff N <= t[b], with
// t[b] = floor(R^2 4^(W - b)) saturated to 2^64 - 1.  False when no R exists.
inline bool beta_thresholds(int n, int64_t k, BetaThresholds *T) {
    const unsigned __int128 R2 = radius2(n, k);
    if (R2 == 0) return false;
    const unsigned __int128 cap = ~(unsigned long long)0;
    for (int b = 0; b <= BETA_MAX; ++b) {
        unsigned __int128 v;
        if (b >= HIST_W) {
            const int sh = 2 * (b - HIST_W);
            v = sh >= 128 ? 0 : R2 >> sh;
        } else {
            const int sh = 2 * (HIST_W - b);
            v = R2 > (cap >> sh) ? cap : R2 << sh;
        }
        T->t[b] = (unsigned long long)(v > cap ? cap : v);
    }
    return true;
}

// The largest beta in [0, 53] with N <= t[beta], or -1 when N exceeds t[0].
__host__ __device__ inline int line_beta(const BetaThresholds &T, unsigned long long N) {
    if (N > T.t[0]) return -1;
    int lo = 0, hi = BETA_MAX;   // t[lo] >= N
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (N <= T.t[mid]) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Group term 4^(W - min(g - e, W)) of one element with exponent e in a group
// whose largest exponent is g >= e.
__host__ __device__ inline unsigned long long group_term(int g, int e) {
    const int d = g - e < HIST_W ? g - e : HIST_W;
    return 1ull << (2 * (HIST_W - d));
}

// The guaranteed exponent for n moduli at depth k: the beta of a line whose
// elements all sit in its top binade (N = k 4^W), which every line gets at
// least; -1 when the moduli cannot hold even that line at beta = 0.
inline int beta_for(int n, int64_t k) {
    BetaThresholds T;
    if (k < 1 || k > ((int64_t)1 << 36) || !beta_thresholds(n, k, &T)) return -1;
    return line_beta(T, (unsigned long long)k << (2 * HIST_W));
}

// Residue of an integer-valued double |x| <= 2^53 modulo MODULI[l], as a
// representative in [-128, 127].  m = 256 takes the low byte of the two's
// complement integer; otherwise q = rint(x / m) is exact to well within 1/2
// (x / m is an integer plus a multiple of 1/m, |x / m| < 2^46), so
// r = x - q m is exact and lies in [-(m+1)/2, (m+1)/2].
__host__ __device__ inline int residue_int8(double x, int l) {
    if (l == 0) return (int)(int8_t)(uint8_t)((uint64_t)(int64_t)x & 0xffu);
    const double m = (double)MODULI[l];
    const double q = rint(x * (1.0 / m));
    int r = (int)fma(-q, m, x);
    if (r > 127) r -= (int)MODULI[l];
    if (r < -128) r += (int)MODULI[l];
    return r;
}

// As residue_int8 for |x| <= 2^51 (beta <= 51).  The quotient estimate
// rint(x * fl(1/m)) then misses rint(x / m) only where |r| would exceed m/2,
// and by at most |x| 2^-52 (1 + 2^-54) / m < 1/2 + 2^-55 in units of m, so
// |r| <= m/2 + 1/2: within int8 for m <= 253, and only r = 128 (m = 255)
// needs folding.  The low byte of x (m = 256) comes from one FP64 add:
// 1.5 * 2^52 + x is exact and its low mantissa bits are x mod 2^32.
__host__ __device__ inline int residue_int8_narrow(double x, int l) {
    if (l == 0) {
        const uint64_t bits = __builtin_bit_cast(uint64_t, x + 6755399441055744.0);
        return (int)(int8_t)(uint8_t)(bits & 0xffu);
    }
    const double m = (double)MODULI[l];
    const double q = rint(x * (1.0 / m));
    int r = (int)fma(-q, m, x);
    if (l == 1 && r > 127) r -= (int)MODULI[1];
    return r;
}

// Direct Chinese remaindering tables for n = 2..16 moduli, in FP64:
//   x = (sum_l r_l W_l) mod P_n,   W_l = (P_n / m_l) ((P_n / m_l)^-1 mod m_l) < P_n,
// with W and P_n in four 40-bit limbs held exactly as doubles (the fourth is
// zero up to 15 moduli), fl(1/P_n), and a threshold just below P_n / 2 for
// the exact range check.
struct CrtTables {
    double w[MAX_MODULI + 1][MAX_MODULI][4];
    double p[MAX_MODULI + 1][4];
    double inv_p[MAX_MODULI + 1];
    double half_p_lo[MAX_MODULI + 1];
};

constexpr int LIMB_BITS = 40;

constexpr CrtTables make_crt_tables() {
    CrtTables T{};
    constexpr unsigned __int128 MASK = ((unsigned __int128)1 << LIMB_BITS) - 1;
    for (int n = MIN_MODULI; n <= MAX_MODULI; ++n) {
        unsigned __int128 P = 1;
        for (int l = 0; l < n; ++l) P *= MODULI[l];
        for (int j = 0; j < 4; ++j) T.p[n][j] = (double)(uint64_t)((P >> (LIMB_BITS * j)) & MASK);
        // 1/P rounded twice (conversion of P, division); only used for an estimate
        T.inv_p[n] = 1.0 / (double)P;
        // P/2 (1 - 2^-36), rounded: below P/2 (1 - 2^-37)
        T.half_p_lo[n] = (double)(P >> 1) * (1.0 - 0x1p-36);
        for (int l = 0; l < n; ++l) {
            const unsigned __int128 Ml = P / MODULI[l];
            const uint32_t y = inverse_mod((uint32_t)(Ml % MODULI[l]), MODULI[l]);
            const unsigned __int128 W = Ml * y;
            for (int j = 0; j < 4; ++j)
                T.w[n][l][j] = (double)(uint64_t)((W >> (LIMB_BITS * j)) & MASK);
        }
    }
    return T;
}

// Limbs a Chinese-remainder value of n moduli needs: P_n < 2^120 for n <= 15
// (P_15 < 2^118), so three 40-bit limbs hold it; P_16 needs four.
__host__ __device__ constexpr int crt_limbs(int n) { return n <= 15 ? 3 : 4; }

// Carries the integer-valued limbs x[0..L-2] into [0, 2^40) (floor division
// by 2^40); the value sum x_j 2^(40 j) is unchanged and the top limb keeps
// its sign.  Exact while every |x_j| stays below 2^53 - 2^14.
template <int L>
__host__ __device__ inline void normalize40(double (&x)[L]) {
#pragma unroll
    for (int j = 0; j + 1 < L; ++j) {
        const double c = floor(x[j] * 0x1p-40);
        x[j] = fma(-c, 0x1p40, x[j]);
        x[j + 1] += c;
    }
}

// Exact integer x with x = r_l (mod m_l), -P/2 < x <= P/2, from residues
// r_l in [0, m_l), l < n, rounded to double (T = make_crt_tables(),
// L = crt_limbs(n); W_l < P_n, so limbs of W at and above L are zero).
//   S_j = sum_l r_l W_lj          exact: every partial sum < 16 * 255 * 2^40 < 2^52
//   q = rint(S / P) from a double estimate (error < 2^-40), so that
//   x = S - q P (limb by limb, exact; after the carries the low limbs lie in
//   [0, 2^40) and the top limb holds the sign) is the symmetric
//   representative unless |x| lies within 2^-39 P of P/2; that case is
//   recognised from the estimate |x| >= the threshold and settled exactly.
// The signed x then goes to double with one rounding: x = a + b + c with
// a = top limbs 2^80 (signed), b = x_1 2^40 and c = x_0 (both in [0, 2^80)
// and [0, 2^40)), each exact.  s + e = a + b exactly (fast two-sum: a = 0 or
// |a| >= 2^80 > b).  e = 0: RN(s + c) is the one rounding.  e != 0: a + b is
// a multiple of 2^40 that needs more than 53 bits, so |a + b| > 2^93, where
// every double and every rounding midpoint is a multiple of 2^40; x lies in
// [a + b, a + b + 2^40) and any c in (0, 2^40) rounds like 2^39.  e + 2^39
// is exact (e is a multiple of 2^40 below 2^67), so RN(s + (e + sticky)) =
// RN(x), for either sign of x.
template <int L>
__host__ __device__ inline double reconstruct_limbs(const CrtTables &T,
                                                    const uint32_t (&r)[MAX_MODULI], int n) {
    static_assert(L == 3 || L == 4, "three or four limbs");
    const double(&W)[MAX_MODULI][4] = T.w[n];
    const double(&P)[4] = T.p[n];
    double x[L];
#pragma unroll
    for (int j = 0; j < L; ++j) x[j] = 0.0;
#pragma unroll
    for (int l = 0; l < MAX_MODULI; ++l) {
        const double rl = l < n ? (double)(r[l] & 0xffu) : 0.0;
#pragma unroll
        for (int j = 0; j < L; ++j) x[j] = fma(rl, W[l][j], x[j]);
    }
    double sd = x[L - 1];
#pragma unroll
    for (int j = L - 2; j >= 0; --j) sd = fma(sd, 0x1p40, x[j]);
    const double q = rint(sd * T.inv_p[n]);   // 0 <= q <= 4080
#pragma unroll
    for (int j = 0; j < L; ++j) x[j] = fma(-q, P[j], x[j]);   // |x_j| < 2^52
    normalize40<L>(x);
    const double hi = L == 4 ? fma(x[L - 1], 0x1p40, x[L - 2]) : x[L - 1];   // |hi| < 2^48
    const double est = fma(fma(hi, 0x1p40, x[1]), 0x1p40, x[0]);
    if (fabs(est) >= T.half_p_lo[n]) {
        // sign of x known from the estimate; t = 2x - sg P, exactly: x leaves
        // (-P/2, P/2] when t > 0 (x > 0) or t <= 0 (x < 0), and moves by -sg P
        const double sg = est < 0.0 ? -1.0 : 1.0;
        double t[L];
#pragma unroll
        for (int j = 0; j < L; ++j) t[j] = fma(2.0, x[j], -sg * P[j]);
        normalize40<L>(t);
        bool zero = true;
#pragma unroll
        for (int j = 0; j < L; ++j) zero = zero && t[j] == 0.0;
        const bool out = sg > 0.0 ? (t[L - 1] >= 0.0 && !zero) : (t[L - 1] < 0.0 || zero);
        if (out) {
#pragma unroll
            for (int j = 0; j < L; ++j) x[j] = fma(-sg, P[j], x[j]);
            normalize40<L>(x);
        }
    }
    const double a = (L == 4 ? fma(x[L - 1], 0x1p40, x[L - 2]) : x[L - 1]) * 0x1p80;
    const double b = x[1] * 0x1p40;
    const double sum = a + b;
    const double e = b - (sum - a);
    const double c = e == 0.0 ? x[0] : e + (x[0] != 0.0 ? 0x1p39 : 0.0);
    return sum + c;
}

__host__ __device__ inline double reconstruct(const CrtTables &T, const uint32_t (&r)[MAX_MODULI],
                                              int n) {
    return crt_limbs(n) == 3 ? reconstruct_limbs<3>(T, r, n) : reconstruct_limbs<4>(T, r, n);
}

}  // namespace modular
}  // namespace mgx
