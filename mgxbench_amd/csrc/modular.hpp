// Modular arithmetic of the INT8 modular split form (Ozaki scheme II):
// the moduli, Garner's mixed-radix tables, residue extraction and the exact
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

// Garner tables for the prefix of the moduli list.  Mixed radix:
//   x = v_0 + v_1 m_0 + v_2 m_0 m_1 + ...,  0 <= v_l < m_l,
//   v_l = ((r_l - sum_{j<l} v_j PM[l][j]) INV[l]) mod m_l,
// with PM[l][j] = (m_0 ... m_{j-1}) mod m_l and INV[l] the inverse of
// (m_0 ... m_{l-1}) mod m_l.  PREFIX[l] = m_0 ... m_{l-1} in four 32-bit limbs.
struct GarnerTables {
    uint32_t pm[MAX_MODULI][MAX_MODULI];
    uint32_t inv[MAX_MODULI];
    uint32_t prefix[MAX_MODULI + 1][4];
};

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

constexpr GarnerTables make_garner_tables() {
    GarnerTables T{};
    for (int l = 0; l < MAX_MODULI; ++l) {
        uint32_t prod = 1 % MODULI[l];
        for (int j = 0; j < MAX_MODULI; ++j) {
            T.pm[l][j] = j < l ? prod : 0;
            if (j < l) prod = (uint32_t)((uint64_t)prod * MODULI[j] % MODULI[l]);
        }
        T.inv[l] = l == 0 ? 1 : inverse_mod(prod, MODULI[l]);
    }
    uint32_t limb[4] = {1, 0, 0, 0};
    for (int l = 0; l <= MAX_MODULI; ++l) {
        for (int w = 0; w < 4; ++w) T.prefix[l][w] = limb[w];
        if (l == MAX_MODULI) break;
        uint64_t carry = 0;
        for (int w = 0; w < 4; ++w) {
            const uint64_t v = (uint64_t)limb[w] * MODULI[l] + carry;
            limb[w] = (uint32_t)v;
            carry = v >> 32;
        }
    }
    return T;
}

constexpr GarnerTables GARNER = make_garner_tables();

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

// Exact integer x with x = r_l (mod m_l), |x| < P/2, from residues
// r_l in [0, m_l), l < n, rounded to double.
__host__ __device__ inline double reconstruct(const uint32_t (&r)[MAX_MODULI], int n) {
    uint32_t v[MAX_MODULI];
    v[0] = r[0];
#pragma unroll
    for (int l = 1; l < MAX_MODULI; ++l) {
        if (l < n) {
            uint32_t s = 0;   // < 15 * 255 * 255 < 2^20
#pragma unroll
            for (int j = 0; j < l; ++j) s += v[j] * GARNER.pm[l][j];
            const uint32_t m = MODULI[l];
            const uint32_t t = s % m;
            const uint32_t d = r[l] >= t ? r[l] - t : r[l] + m - t;
            v[l] = d * GARNER.inv[l] % m;
        }
    }
    // x = sum_l v_l PREFIX[l] in 32-bit limbs (each limb sum < 16 * 2^40)
    uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < MAX_MODULI; ++l) {
        if (l < n) {
#pragma unroll
            for (int w = 0; w < 4; ++w) acc[w] += (uint64_t)v[l] * GARNER.prefix[l][w];
        }
    }
    uint32_t x[4];
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint64_t t = acc[w] + carry;
        x[w] = (uint32_t)t;
        carry = t >> 32;
    }
    // P_n and the half-range test x >= P - x  <=>  x > P/2 (P_n is even: m_0 = 256)
    uint32_t P[4] = {0, 0, 0, 0};
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t pw = 0;
#pragma unroll
        for (int l = 1; l <= MAX_MODULI; ++l)
            if (l == n) pw = GARNER.prefix[l][w];
        P[w] = pw;
    }
    // y = P - x
    uint32_t y[4];
    int64_t borrow = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int64_t t = (int64_t)P[w] - (int64_t)x[w] - borrow;
        y[w] = (uint32_t)t;
        borrow = t < 0 ? 1 : 0;
    }
    // x > P/2  <=>  x > y
    bool greater = false, decided = false;
#pragma unroll
    for (int w = 3; w >= 0; --w) {
        if (!decided && x[w] != y[w]) {
            greater = x[w] > y[w];
            decided = true;
        }
    }
    uint32_t mag[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) mag[w] = greater ? y[w] : x[w];
    const double d = u128_to_double_rn(((uint64_t)mag[3] << 32) | mag[2],
                                       ((uint64_t)mag[1] << 32) | mag[0]);
    return greater ? -d : d;
}

}  // namespace modular
}  // namespace mgx
