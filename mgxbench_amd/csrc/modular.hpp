// Modular arithmetic of the INT8 modular split form (Ozaki scheme II):
// the moduli, the Chinese-remainder tables, residue extraction and the exact
// reconstruction of each output.  Internal to libmgx.
//
// Published algorithm: K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique" (2025).  A row of A is scaled to integers
// A' = rint(A 2^(beta-e_i)), |A'| <= 2^beta, B per column likewise; for each
// modulus m_l the INT8 matrix cores form (A' mod m_l)(B' mod m_l) exactly in
// INT32, reduced mod m_l; Chinese remaindering over the n moduli recovers the
// integer C' = A'B' exactly because 2 k 4^beta < P = m_0 ... m_{n-1}.  The
// output is ldexp(RN(C'), e_i + f_j - 2 beta): one rounding.
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

// Scaling exponent for n moduli and depth k: the largest beta in [0, 53] with
// 2 k 4^beta < P_n, or -1 when even beta = 0 does not fit.
inline int beta_for(int n, int64_t k) {
    if (n < MIN_MODULI || n > MAX_MODULI || k < 1) return -1;
    unsigned __int128 P = 1;
    for (int l = 0; l < n; ++l) P *= MODULI[l];
    unsigned __int128 lhs = (unsigned __int128)2 * (uint64_t)k;
    if (lhs >= P) return -1;
    int beta = 0;
    while (beta < 53 && lhs <= (P - 1) / 4 && lhs * 4 < P) {
        lhs *= 4;
        ++beta;
    }
    return beta;
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

// Correctly rounded (to nearest, ties to even) conversion of the 128-bit
// magnitude hi:lo to double.
__host__ __device__ inline double u128_to_double_rn(uint64_t hi, uint64_t lo) {
    if (hi == 0 && lo == 0) return 0.0;
    uint64_t top, rest;
    int e;
    if (hi != 0) {
        const int s = __builtin_clzll(hi);
        top = s ? (hi << s) | (lo >> (64 - s)) : hi;
        rest = s ? lo << s : lo;
        e = 75 - s;   // value = (top 2^64 + rest) 2^-s,  top = mant 2^11 + round bits
    } else {
        const int s = __builtin_clzll(lo);
        top = lo << s;
        rest = 0;
        e = 11 - s;
    }
    uint64_t mant = top >> 11;
    const uint32_t rb = (uint32_t)(top & 0x7ffu);
    if (rb > 0x400u || (rb == 0x400u && (rest != 0 || (mant & 1u)))) ++mant;   // mant <= 2^53
    return ldexp((double)mant, e);
}

// Direct Chinese remaindering tables for n = 2..16 moduli:
//   x = (sum_l r_l W_l) mod P_n,   W_l = (P_n / m_l) ((P_n / m_l)^-1 mod m_l) < P_n,
// with W and P_n in eight 16-bit limbs and 1/P_n as a double.
struct CrtTables {
    uint32_t w[MAX_MODULI + 1][MAX_MODULI][8];
    uint32_t p[MAX_MODULI + 1][8];
    double inv_p[MAX_MODULI + 1];
};

constexpr CrtTables make_crt_tables() {
    CrtTables T{};
    for (int n = MIN_MODULI; n <= MAX_MODULI; ++n) {
        unsigned __int128 P = 1;
        for (int l = 0; l < n; ++l) P *= MODULI[l];
        for (int w = 0; w < 8; ++w) T.p[n][w] = (uint32_t)(P >> (16 * w)) & 0xffffu;
        // 1/P rounded twice (conversion of P, division); only used for an estimate
        T.inv_p[n] = 1.0 / (double)P;
        for (int l = 0; l < n; ++l) {
            const unsigned __int128 Ml = P / MODULI[l];
            const uint32_t y = inverse_mod((uint32_t)(Ml % MODULI[l]), MODULI[l]);
            const unsigned __int128 W = Ml * y;
            for (int w = 0; w < 8; ++w) T.w[n][l][w] = (uint32_t)(W >> (16 * w)) & 0xffffu;
        }
    }
    return T;
}

// Exact integer x with x = r_l (mod m_l), |x| < P/2, from residues
// r_l in [0, m_l), l < n, rounded to double (T = make_crt_tables()).
//   S = sum_l r_l W_l       (16-bit limbs, each limb sum < 16 * 2^24)
//   q = floor(S / P) from a double estimate, corrected by one either way
//   x = S - q P in [0, P);  C' = x, or x - P when x > P / 2.
__host__ __device__ inline double reconstruct(const CrtTables &T, const uint32_t (&r)[MAX_MODULI],
                                              int n) {
    const uint32_t(&W)[MAX_MODULI][8] = T.w[n];
    const uint32_t(&P)[8] = T.p[n];
    uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < MAX_MODULI; ++l) {
        const uint32_t rl = l < n ? (r[l] & 0xffu) : 0u;
#pragma unroll
        for (int w = 0; w < 8; ++w) acc[w] += rl * (W[l][w] & 0xffffu);
    }
    // normalise to 16-bit limbs s[0..8] (S < 16 * 256 * P < 2^139)
    uint32_t sl[9];
    uint32_t carry = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const uint32_t t = acc[w] + carry;
        sl[w] = t & 0xffffu;
        carry = t >> 16;
    }
    sl[8] = carry;
    double sd = 0.0;
#pragma unroll
    for (int w = 8; w >= 0; --w) sd = sd * 65536.0 + (double)sl[w];
    int32_t q = (int32_t)floor(sd * T.inv_p[n]);   // 0 <= q < 4096, off by at most one
    // x = S - q P (signed 16-bit limb arithmetic with borrow)
    int32_t xl[9];
    int32_t borrow = 0;
#pragma unroll
    for (int w = 0; w < 9; ++w) {
        const int32_t pw = w < 8 ? (int32_t)(P[w] & 0xffffu) : 0;
        int32_t t = (int32_t)sl[w] - q * pw - borrow;   // |q * pw| < 2^28
        borrow = 0;
        // floor division of t by 2^16 (arithmetic shift) -> next borrow
        borrow = -(t >> 16);
        xl[w] = t & 0xffff;
    }
    // the value is sum xl[w] 2^(16w) - borrow 2^144; borrow != 0 means x < 0
    auto add_p = [&](int sign) {
        int32_t c = 0;
#pragma unroll
        for (int w = 0; w < 9; ++w) {
            const int32_t pw = w < 8 ? (int32_t)(P[w] & 0xffffu) : 0;
            const int32_t t = xl[w] + sign * pw + c;
            c = t >> 16;
            xl[w] = t & 0xffff;
        }
        return c;
    };
    if (borrow != 0) borrow -= add_p(1);   // q was one too large
    // x >= P ?  compare limbs from the top
    {
        bool ge = true, decided = false;
#pragma unroll
        for (int w = 8; w >= 0; --w) {
            const uint32_t pw = w < 8 ? (P[w] & 0xffffu) : 0u;
            if (!decided && (uint32_t)xl[w] != pw) {
                ge = (uint32_t)xl[w] > pw;
                decided = true;
            }
        }
        if (ge) add_p(-1);   // q was one too small
    }
    // symmetric range: x > P/2  <=>  2x > P
    uint32_t x32[4] = {(uint32_t)xl[0] | ((uint32_t)xl[1] << 16),
                       (uint32_t)xl[2] | ((uint32_t)xl[3] << 16),
                       (uint32_t)xl[4] | ((uint32_t)xl[5] << 16),
                       (uint32_t)xl[6] | ((uint32_t)xl[7] << 16)};
    uint32_t p32[4] = {(P[0] & 0xffffu) | (P[1] << 16), (P[2] & 0xffffu) | (P[3] << 16),
                       (P[4] & 0xffffu) | (P[5] << 16), (P[6] & 0xffffu) | (P[7] << 16)};
    uint32_t y[4];
    int64_t bo = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int64_t t = (int64_t)p32[w] - (int64_t)x32[w] - bo;
        y[w] = (uint32_t)t;
        bo = t < 0 ? 1 : 0;
    }
    bool greater = false, decided = false;
#pragma unroll
    for (int w = 3; w >= 0; --w) {
        if (!decided && x32[w] != y[w]) {
            greater = x32[w] > y[w];
            decided = true;
        }
    }
    uint32_t mag[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) mag[w] = greater ? y[w] : x32[w];
    const double d = u128_to_double_rn(((uint64_t)mag[3] << 32) | mag[2],
                                       ((uint64_t)mag[1] << 32) | mag[0]);
    return greater ? -d : d;
}

}  // namespace modular
}  // namespace mgx
