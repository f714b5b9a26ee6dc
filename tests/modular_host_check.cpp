// Host build of the modular form's arithmetic (mgxbench_amd/csrc/modular.hpp)
// checked against exact 128-bit integer arithmetic: the Chinese-remainder
// reconstruction (symmetric representative, one correct rounding) and the
// residue extraction.  Built with g++ by tests/test_modular_host.py; prints
// one summary line and exits non-zero on any mismatch.
#include <math.h>
#include <stdint.h>
#define __host__
#define __device__
#include "modular.hpp"

#include <cstdio>
#include <random>

using namespace mgx::modular;
typedef __int128 i128;
typedef unsigned __int128 u128;

static const CrtTables TAB = make_crt_tables();
static long checked = 0, bad = 0;

static u128 product(int n) {
    u128 P = 1;
    for (int l = 0; l < n; ++l) P *= MODULI[l];
    return P;
}

// reconstruct() of the residues of x against the correctly rounded x
// (g++ converts __int128 to double to nearest, ties to even)
static void check_value(int n, i128 x, const char *what) {
    uint32_t r[MAX_MODULI] = {};
    for (int l = 0; l < n; ++l) {
        i128 t = x % (i128)MODULI[l];
        if (t < 0) t += MODULI[l];
        r[l] = (uint32_t)t;
    }
    const double got = reconstruct(TAB, r, n), want = (double)x;
    ++checked;
    if (got != want || signbit(got) != signbit(want)) {
        if (bad < 20) printf("%s n=%d: got %.17g want %.17g\n", what, n, got, want);
        ++bad;
    }
}

int main() {
    std::mt19937_64 g(20261203);
    for (int n = MIN_MODULI; n <= MAX_MODULI; ++n) {
        const u128 P = product(n);
        const i128 half = (i128)(P / 2);   // representatives: (-P/2, P/2]
        // random magnitudes at every scale
        for (int it = 0; it < 60000; ++it) {
            u128 u = ((u128)g() << 64) | g();
            const int bits = (int)(g() % 128);
            u = bits == 0 ? 0 : u >> (128 - bits);
            i128 x = (i128)(u % (u128)half);
            if (g() & 1) x = -x;
            check_value(n, x, "random");
        }
        // next to +-P/2, where the quotient estimate can miss by one
        for (int it = 0; it < 20000; ++it) {
            const i128 d = it < 1000 ? it : (i128)(g() >> (g() % 64));
            i128 x = half - d;
            if (x <= -half) continue;
            if (g() & 1) x = -x;
            if (x == -half) x = half;
            check_value(n, x, "near P/2");
        }
        // small values of both signs
        for (i128 x = -300; x <= 300; ++x)
            if (x > -half && x <= half) check_value(n, x, "small");
    }
    // every representative for two moduli
    {
        const int64_t P = (int64_t)product(2);
        for (int64_t x = -P / 2 + 1; x <= P / 2; ++x) check_value(2, x, "two moduli");
    }
    // exact rounding ties and near-ties at every scale up to 2^117
    for (int n = 14; n <= MAX_MODULI; ++n) {
        const i128 half = (i128)(product(n) / 2);
        for (int E = 1; E <= 64; ++E) {
            for (int it = 0; it < 100; ++it) {
                const i128 M = (i128)((g() >> 11) | (1ull << 52));
                const i128 tie = (M << E) + ((i128)1 << (E - 1));
                const i128 lim = ((i128)1 << (E - 1)) - 1;
                const i128 offs[] = {0, 1, -1, (i128)1 << 39, -((i128)1 << 39),
                                     ((i128)1 << 40) - 1, -(((i128)1 << 40) - 1), lim, -lim};
                for (i128 o : offs) {
                    if (o > lim || o < -lim || tie + o >= half) continue;
                    check_value(n, tie + o, "tie");
                    check_value(n, -(tie + o), "tie");
                }
            }
        }
    }
    // residues: general form for |x| <= 2^53, narrow form for |x| <= 2^51
    long residues = 0;
    auto check_residues = [&](int64_t xi, bool narrow) {
        for (int l = 0; l < MAX_MODULI; ++l) {
            const int a = residue_int8((double)xi, l);
            const int b = narrow ? residue_int8_narrow((double)xi, l) : a;
            ++residues;
            if (a != b || a < -128 || a > 127 || (xi - a) % (int64_t)MODULI[l] != 0) {
                if (bad < 40) printf("residue x=%lld l=%d: %d %d\n", (long long)xi, l, a, b);
                ++bad;
            }
        }
    };
    const int64_t LIM51 = (int64_t)1 << 51, LIM53 = (int64_t)1 << 53;
    for (int64_t xi = -5000; xi <= 5000; ++xi) check_residues(xi, true);
    for (int64_t d = 0; d <= 3; ++d) {
        check_residues(LIM51 - d, true);
        check_residues(-LIM51 + d, true);
        check_residues(LIM53 - d, false);
        check_residues(-LIM53 + d, false);
    }
    for (int it = 0; it < 200000; ++it) {
        int64_t xi = (int64_t)(g() % (uint64_t)(LIM51 + 1)) >> (g() % 52);
        if (g() & 1) xi = -xi;
        check_residues(xi, true);
        int64_t xw = (int64_t)(g() % (uint64_t)(LIM53 + 1)) >> (g() % 54);
        if (g() & 1) xw = -xw;
        check_residues(xw, false);
    }
    // narrow form next to the half-way points of the odd moduli near 2^51
    for (int l = 1; l < MAX_MODULI; ++l) {
        const int64_t m = MODULI[l];
        for (int it = 0; it < 2000; ++it) {
            const int64_t j = LIM51 / m - (int64_t)(g() % 100000);
            for (int64_t off = -2; off <= 2; ++off) {
                const int64_t xi = j * m + m / 2 + off;
                if (xi <= LIM51) {
                    check_residues(xi, true);
                    check_residues(-xi, true);
                }
            }
        }
    }
    // per-line exponents: the threshold search against the definition
    // (largest b in [0, 53] with 4^b N <= R^2 4^W) evaluated by exact division
    long lines = 0;
    const int64_t depths[] = {1, 7, 64, 1000, 4096, 16384, 32768, 131072};
    for (int n = MIN_MODULI; n <= MAX_MODULI; ++n) {
        for (int64_t k : depths) {
            BetaThresholds T;
            const u128 R2 = radius2(n, k);
            if (!beta_thresholds(n, k, &T)) {
                if (R2 != 0) { printf("thresholds refused n=%d k=%lld\n", n, (long long)k); ++bad; }
                continue;
            }
            const unsigned long long nmax = (unsigned long long)k << (2 * HIST_W);
            for (int it = 0; it < 4000; ++it) {
                unsigned long long N = it == 0 ? nmax : it == 1 ? 0 : g() % (nmax + 1);
                int want = -1;
                for (int b = 0; b <= BETA_MAX; ++b) {
                    const bool ok = b >= HIST_W ? (u128)N <= (R2 >> (2 * (b - HIST_W)))
                                                : (((u128)N + ((u128)1 << (2 * (HIST_W - b))) - 1)
                                                   >> (2 * (HIST_W - b))) <= R2;
                    if (ok) want = b;
                }
                const int got = line_beta(T, N);
                ++lines;
                if (got != want) {
                    if (bad < 20) printf("line_beta n=%d k=%lld N=%llu: %d want %d\n", n,
                                         (long long)k, N, got, want);
                    ++bad;
                }
                if (it == 0 && got != beta_for(n, k)) {
                    printf("beta_for n=%d k=%lld: %d vs dense line %d\n", n, (long long)k,
                           beta_for(n, k), got);
                    ++bad;
                }
            }
        }
    }
    printf("reconstructions %ld residues %ld line exponents %ld bad %ld\n", checked, residues,
           lines, bad);
    return bad != 0;
}
