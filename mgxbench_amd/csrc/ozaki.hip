// FP64-accurate split-precision GEMM on the INT8 matrix cores of gfx950, in
// two forms that share one INT8 product kernel, plus the digit-slice form's
// bf16 variant (the same digits on the bf16 matrix cores, same result).
//
// No counterpart in the reference: both compute the same local product as
// ABMultiply / ABMultiplyAccumulateAsync (1D-row-block/mulmat_kernel.cu:86,
// 148) through a different arithmetic.
//
// Common first step: e_i = frexp exponent of max_k |A_ik| (so |A_ik| < 2^e_i),
// f_j likewise for column j of B.  A row/column holding Inf/NaN gets a
// sentinel and its outputs become NaN.  Operands are stored k-tile blocked,
// int8, B per column (n-major, k contiguous).
//
// Digit-slice form (MGX_FORM_INT8, Ozaki scheme I):
//  1. x = A_ik 2^-e_i in (-1,1) is cut into s signed digits by round to
//     nearest: sigma_1 = rint(x 2^6), then sigma_p = rint(256 r) of the
//     previous remainder r, with a carry pass that keeps every digit in
//     [-128, 127] (slice_digits).  Place values are 2^(2-8p); the residual
//     after s slices is at most 2^(1-8s).  Every step is exact in FP64.
//  2. For each weight group t = 2..s+1 (slice pairs p+q = t, place value
//     2^(4-8t)), the INT8 MFMA accumulates G_t = sum_{p+q=t} sum_k sigmaA_p
//     sigmaB_q exactly in INT32 over K chunks with
//     pairs(t) * chunk * 64 * 128^2 < 2^31.
//  3. After each chunk the sums are folded into C in FP64:
//     x = ldexp((double)G, e_i + f_j + 4 - 8t) (exact); the first fold writes
//     C = x (or C_in + x for the accumulate form), later folds C = C + x.
//
// Modular form (MGX_FORM_INT8_MOD, Ozaki scheme II; modular.hpp):
//  1. A'_ik = rint(A_ik 2^(beta_i-e_i)) (|A'| <= 2^beta_i), B' likewise per
//     column with delta_j; each line's exponent comes from an exponent-group
//     sum of the line by a Cauchy-Schwarz bound that keeps every |C'_ij|
//     below P/2, P the product of the first n moduli (modular.hpp).
//     Operands are the residues of A', B' modulo each m_l, as int8.
//  2. For each modulus the INT8 MFMA forms the residue product exactly in
//     INT32 (reduced mod m_l between K chunks of at most 2046 tiles) and
//     stores it mod m_l as one byte per output.
//  3. A reconstruction kernel recovers the integer C' = A'B' from its n
//     residues (direct Chinese remaindering in FP64 on exact 40-bit limbs),
//     rounds it once to FP64 and writes
//     C = ldexp(C', e_i - beta_i + f_j - delta_j) (or C_in + that).
// The modular form needs n products instead of s(s+1)/2 and its only error
// is the quantisation of each line of A and B to its beta bits.
#include "mgx_internal.hpp"
#include "modular.hpp"

#include <algorithm>
#include <type_traits>

namespace {

using mgx::modular::MAX_MODULI;
using mgx::modular::MIN_MODULI;
using mgx::modular::MODULI;

constexpr int KT = 64;             // k depth of one INT8 MFMA step / LDS stage
constexpr int K_ALIGN = 256;       // operand rows are zero-padded to this multiple
constexpr int EXP_BAD = 0x40000000;
constexpr int MIN_SLICES = 1;
constexpr int MAX_SLICES = 9;
constexpr long long SLICE_MAX_PRODUCT = 128ll * 128ll;  // largest |sigma_p sigma_q|
constexpr int MOD_CHUNK_TILES = 2046;   // 2046 * 64 * 128^2 + 256 < 2^31

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Chinese-remainder tables of the modular form (n = 2..16), in constant memory.
__constant__ mgx::modular::CrtTables CRT_TABLES = mgx::modular::make_crt_tables();

// ---- exponent scan --------------------------------------------------------

__global__ void zero_u64_kernel(unsigned long long *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__device__ inline unsigned long long abs_bits(double x) {
    return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
}

// Per-element exponent data: e = the frexp exponent of a finite nonzero x
// (|x| < 2^e, subnormals included; v_frexp_exp_i32_f64) or NO_EXP, and the
// line-maximum code: 0 for zero, e + CODE_BIAS for a finite nonzero x,
// CODE_BAD for Inf / NaN.  A line's largest code is its maximum's code
// (frexp exponents order like magnitudes), taken with integer atomicMax.
constexpr int NO_EXP = -(1 << 20);
constexpr int CODE_BIAS = 2048;
constexpr unsigned long long CODE_BAD = 0x7fffffffull;
constexpr int CLASS_FINITE_NONZERO = 0x198;   // -normal, -denormal, +denormal, +normal
constexpr int CLASS_INF_NAN = 0x207;          // signalling / quiet NaN, -Inf, +Inf
__device__ inline void elem_exp(double x, int &e, unsigned &code) {
    const int fe = __builtin_amdgcn_frexp_exp(x);
    const bool ok = __builtin_amdgcn_class(x, CLASS_FINITE_NONZERO);
    const bool bad = __builtin_amdgcn_class(x, CLASS_INF_NAN);
    e = ok ? fe : NO_EXP;
    code = ok ? (unsigned)(fe + CODE_BIAS) : bad ? (unsigned)CODE_BAD : 0u;
}

// Group term of mgx::modular's exponent rule, 0 for an element without an
// exponent (the group maximum g is at least e).
__device__ inline unsigned group_term32(int g, int e) {
    const int d = min(g - e, mgx::modular::HIST_W);
    return e != NO_EXP ? 1u << (2 * (mgx::modular::HIST_W - d)) : 0u;
}

// Row scan of A (m x k, lda): the row-maximum codes and, for the modular form
// (SUMS), the exponent-group sums N of the per-line scaling rule
// (mgx::modular, groups of 16 consecutive k).  grid.x covers k in segments of
// 256*16, grid.y strides over rows.  For a fixed i the 64 lanes of a wave
// hold 64 consecutive k, so a group is one 16-lane DPP row, whose maximum
// takes four DPP steps (quad swaps, then the half-row and row mirrors; every
// lane of the wave is active).  Codes and sums
// reach memory through integer atomicMax / atomicAdd, so the result does not
// depend on the order of the workgroups.
template <bool SUMS>
__global__ __launch_bounds__(256) void row_scan_kernel(const double *__restrict__ A, int64_t m,
                                                       int64_t k, int64_t lda,
                                                       unsigned long long *rowmax,
                                                       unsigned long long *rowsum) {
    __shared__ unsigned long long sh[4], ss[4];
    const int64_t kbase = (int64_t)blockIdx.x * (256 * 16);
    for (int64_t r = blockIdx.y; r < m; r += gridDim.y) {
        const double *row = A + r * lda + kbase + threadIdx.x;
        double v[16];
        if (kbase + 256 * 16 <= k) {
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = __builtin_nontemporal_load(&row[i * 256]);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                v[i] = kbase + i * 256 + threadIdx.x < k ? __builtin_nontemporal_load(&row[i * 256])
                                                         : 0.0;
        }
        unsigned mx = 0, ns = 0;   // ns <= 16 * 4^12 = 2^28
        int e[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            unsigned code;
            elem_exp(v[i], e[i], code);
            mx = max(mx, code);
        }
        if constexpr (SUMS) {
            // the 16 group maxima stage by stage, so consecutive DPP steps
            // work on different values
            int g[16];
#pragma unroll
            for (int i = 0; i < 16; ++i)
                g[i] = max(e[i], __builtin_amdgcn_update_dpp(e[i], e[i], 0xB1, 0xF, 0xF, true));
#pragma unroll
            for (int i = 0; i < 16; ++i)
                g[i] = max(g[i], __builtin_amdgcn_update_dpp(g[i], g[i], 0x4E, 0xF, 0xF, true));
#pragma unroll
            for (int i = 0; i < 16; ++i)
                g[i] = max(g[i], __builtin_amdgcn_update_dpp(g[i], g[i], 0x141, 0xF, 0xF, true));
#pragma unroll
            for (int i = 0; i < 16; ++i)
                g[i] = max(g[i], __builtin_amdgcn_update_dpp(g[i], g[i], 0x140, 0xF, 0xF, true));
#pragma unroll
            for (int i = 0; i < 16; ++i) ns += group_term32(g[i], e[i]);
        }
        unsigned long long mx64 = mx, ns64 = ns;
        for (int off = 32; off > 0; off >>= 1) {
            mx64 = max(mx64, (unsigned long long)__shfl_xor(mx64, off));
            if constexpr (SUMS) ns64 += (unsigned long long)__shfl_xor(ns64, off);
        }
        if ((threadIdx.x & 63) == 0) {
            sh[threadIdx.x >> 6] = mx64;
            ss[threadIdx.x >> 6] = ns64;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            mx64 = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
            if (mx64) atomicMax(rowmax + r, mx64);
            if constexpr (SUMS) {
                ns64 = ss[0] + ss[1] + ss[2] + ss[3];
                if (ns64) atomicAdd(rowsum + r, ns64);
            }
        }
        __syncthreads();
    }
}

// Column scan of B (k x n, ldb): one lane per column, grid.y splits k in
// segments of `seg` rows (a multiple of 16, chosen by the host so that short
// depths still give a full grid), read 16 rows (one group) at a time.
template <bool SUMS>
__global__ __launch_bounds__(256) void col_scan_kernel(const double *__restrict__ B, int64_t k,
                                                       int64_t n, int64_t ldb, int64_t seg,
                                                       unsigned long long *colmax,
                                                       unsigned long long *colsum) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int64_t k0 = (int64_t)blockIdx.y * seg;
    const int64_t k1 = min(k, k0 + seg);
    unsigned mx = 0;
    unsigned long long ns = 0;
    for (int64_t kq = k0; kq < k1; kq += 16) {
        double v[16];
        const double *col = B + kq * ldb + c;
        // read-once stream: non-temporal loads
        if (kq + 16 <= k1) {
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = __builtin_nontemporal_load(&col[i * ldb]);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                v[i] = kq + i < k1 ? __builtin_nontemporal_load(&col[i * ldb]) : 0.0;
        }
        int e[16], g = NO_EXP;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            unsigned code;
            elem_exp(v[i], e[i], code);
            mx = max(mx, code);
            g = max(g, e[i]);
        }
        if constexpr (SUMS) {
            unsigned gs = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) gs += group_term32(g, e[i]);
            ns += gs;
        }
    }
    if (mx) atomicMax(colmax + c, (unsigned long long)mx);
    if constexpr (SUMS)
        if (ns) atomicAdd(colsum + c, ns);
}

// Line exponent from a line-maximum code (0 for an all-zero line).
__device__ inline int exponent_of_code(unsigned long long code) {
    if (code == 0) return 0;
    if (code >= CODE_BAD) return EXP_BAD;
    return (int)code - CODE_BIAS;
}

// Line exponents e from the maximum codes; for the modular form (MOD) also
// the line's scale shift e - beta, beta the line's scaling exponent from its
// group sum (mgx::modular::line_beta); EXP_BAD for a non-finite line.  The
// scaled line is X 2^-shift and the product's exponent is the sum of the two
// lines' shifts.
template <bool MOD>
__global__ void line_scales_kernel(const unsigned long long *maxcode,
                                   const unsigned long long *sums, int32_t *e, int32_t *shift,
                                   int64_t n, mgx::modular::BetaThresholds T) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ei = exponent_of_code(maxcode[i]);
        e[i] = ei;
        if constexpr (MOD)
            shift[i] = ei == EXP_BAD ? EXP_BAD : ei - mgx::modular::line_beta(T, sums[i]);
    }
}

// ---- slice extraction -----------------------------------------------------

// Slice digits of one scaled value x in (-1,1), most significant first:
//   y = x * 2^6, sigma_1 = rint(y), r = y - sigma_1            (|sigma_1| <= 64)
//   y = 256 r,   sigma_p = rint(y), r = y - sigma_p  (p >= 2)  (|sigma_p| <= 128)
// then a carry pass from the least significant digit up turns every 128 into
// -128 and adds 1 to the digit above (|sigma_1| <= 65 afterwards), so all
// digits fit int8.  Every step is exact in FP64, and
//   x = sum_p sigma_p 2^(2-8p) + r_s 2^(2-8s),  |r_s| <= 1/2.
__device__ inline void slice_digits(double x, int s, int *dig) {
    double y = x * 64.0;
    for (int p = 0; p < s; ++p) {
        const double d = rint(y);
        dig[p] = (int)d;
        y = (y - d) * 256.0;
    }
    for (int p = s - 1; p > 0; --p) {
        if (dig[p] >= 128) {
            dig[p] -= 256;
            dig[p - 1] += 1;
        }
    }
}

// Writes the s slices of 16 consecutive scaled values x[0..15] as 16 bytes
// each.  dst points at slice 1; consecutive slices are `slice_stride` apart.
__device__ inline void emit_slices16(const double *x, int s, int8_t *dst, int64_t slice_stride) {
    unsigned int words[MAX_SLICES][4];
#pragma unroll
    for (int p = 0; p < MAX_SLICES; ++p)
#pragma unroll
        for (int w = 0; w < 4; ++w) words[p][w] = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int dig[MAX_SLICES];
        slice_digits(x[i], s, dig);
#pragma unroll
        for (int p = 0; p < MAX_SLICES; ++p)
            if (p < s) words[p][i / 4] |= ((unsigned int)(dig[p] & 0xff)) << (8 * (i % 4));
    }
#pragma unroll
    for (int p = 0; p < MAX_SLICES; ++p) {
        if (p < s) {
            const v4i packed = {(int)words[p][0], (int)words[p][1], (int)words[p][2],
                                (int)words[p][3]};
            *(v4i *)(dst + (int64_t)p * slice_stride) = packed;
        }
    }
}

// BF16 digit planes: the same digits as bf16 values (every integer of
// magnitude <= 256 is exact in bf16's 8 significant bits), 16 consecutive
// values as 32 bytes.  dst points at slice 1; consecutive slices are
// `slice_stride` apart.
__device__ inline void emit_slices16_bf16(const double *x, int s, int8_t *dst,
                                          int64_t slice_stride) {
    unsigned int words[MAX_SLICES][8];
#pragma unroll
    for (int p = 0; p < MAX_SLICES; ++p)
#pragma unroll
        for (int w = 0; w < 8; ++w) words[p][w] = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int dig[MAX_SLICES];
        slice_digits(x[i], s, dig);
#pragma unroll
        for (int p = 0; p < MAX_SLICES; ++p)
            if (p < s) {
                // bf16 = the upper half of the (exact) fp32 pattern
                const unsigned int h = __float_as_uint((float)dig[p]) >> 16;
                words[p][i / 2] |= h << (16 * (i % 2));
            }
    }
#pragma unroll
    for (int p = 0; p < MAX_SLICES; ++p) {
        if (p < s) {
            v4i *d = (v4i *)(dst + (int64_t)p * slice_stride);
            d[0] = v4i{(int)words[p][0], (int)words[p][1], (int)words[p][2], (int)words[p][3]};
            d[1] = v4i{(int)words[p][4], (int)words[p][5], (int)words[p][6], (int)words[p][7]};
        }
    }
}

// Writes the n residues (modular form) of 16 consecutive integer-valued
// scaled values x[0..15] as 16 bytes each.  dst points at modulus 0;
// consecutive moduli are `plane_stride` apart.
template <bool NARROW>
__device__ inline void emit_residues16(const double *x, int n, int8_t *dst, int64_t plane_stride) {
#pragma unroll
    for (int l = 0; l < MAX_MODULI; ++l) {
        if (l < n) {
            unsigned int words[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int r = NARROW ? mgx::modular::residue_int8_narrow(x[i], l)
                                     : mgx::modular::residue_int8(x[i], l);
                words[i / 4] |= ((unsigned int)(r & 0xff)) << (8 * (i % 4));
            }
            const v4i packed = {(int)words[0], (int)words[1], (int)words[2], (int)words[3]};
            *(v4i *)(dst + (int64_t)l * plane_stride) = packed;
        }
    }
}

// ---- operand layout ------------------------------------------------------------
//
// Operands are stored k-tile blocked: plane p (slice p or modulus p) of A is
// KTC = Kp/64 blocks, block kt holding rows 0..Mp-1 (Mp = M rounded up to 256,
// padded rows zero) x 64 bytes of k:
//     Aop[((p * KTC + kt) * Mp + row) * 64 + (k % 64)].
// B is stored the same way per column of B (n-major, k contiguous).  A 16-row
// x 64-byte piece of a block is therefore one contiguous KiB in HBM.
constexpr int ROW_ALIGN = 256;

// Per-element operand value: scheme I scales into (-1,1), scheme II to an
// integer of at most beta bits (beta = the line's scaling exponent).
template <bool MOD>
__device__ inline double scaled_operand(double a, int e, int beta) {
    return MOD ? rint(ldexp(a, beta - e)) : ldexp(a, -e);
}

// One workgroup converts a 64 x 64 tile of operand lines (rows of A or
// columns of B) by 64 k.  The tile is first staged in LDS by coalesced loads
// (a wave reads one 512-byte row of the source matrix per instruction: a row
// of A is a line, a row of B crosses 64 lines), then lane t of wave w
// converts line 16 w + t % 16, k values 16 (t / 16) .. +15, and writes 16
// bytes per plane; the 64 lanes of a wave write one contiguous KiB per plane.
// LDS rows are padded to 528 bytes so that both read patterns are free of
// bank conflicts.
//
// BF (digit slices only): bf16 planes, 2 bytes per value, so a 64-byte plane
// row holds 32 k values and the plane has KTC = 2 Kp / 64 row blocks; the 16
// values of a lane are 32 contiguous bytes of one row.
template <bool MOD, bool COLS, bool BF = false>
__global__ __launch_bounds__(256) void operand_tile_kernel(const double *__restrict__ X,
                                                           int64_t lines, int64_t k, int64_t ld,
                                                           const int32_t *__restrict__ ex,
                                                           const int32_t *__restrict__ lshift,
                                                           int planes, int64_t Lp, int64_t KTC,
                                                           int8_t *__restrict__ Xop) {
    constexpr int STR = 66;   // doubles per padded LDS row
    __shared__ double tile[64 * STR];
    // Consecutive workgroups take consecutive 512-byte segments of the same
    // 64 source rows (k tiles for A, line blocks for B), so DRAM pages are
    // read in one pass.
    const int64_t kt = COLS ? blockIdx.y : blockIdx.x;
    const int64_t l0 = (int64_t)(COLS ? blockIdx.x : blockIdx.y) * 64;
    const int64_t k0 = kt * KT;
    const int t = threadIdx.x;
    // Stage the 64 x 64 source tile (source rows are lines for A, k for B).
    // Interior tiles of a 16-byte aligned matrix with an even leading
    // dimension take 16-byte loads: element pair e = 2 (t + 256 i) is (source
    // row e / 64, source columns e % 64, +1); the source is read once, so
    // the loads are non-temporal.  Other tiles load element by element.
    const bool full = l0 + 64 <= lines && k0 + 64 <= k;
    const bool pairs = ((reinterpret_cast<uintptr_t>(X) & 15) == 0) && ((ld & 1) == 0);
    if (full && pairs) {
        typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = 2 * (t + 256 * i);
            const int sr = e >> 6, sc = e & 63;
            const d2 *src = (const d2 *)(COLS ? &X[(k0 + sr) * ld + l0 + sc]
                                              : &X[(l0 + sr) * ld + k0 + sc]);
            *(d2 *)&tile[sr * STR + sc] = __builtin_nontemporal_load(src);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = t + 256 * i;
            const int sr = e >> 6, sc = e & 63;
            double v = 0.0;
            if (!COLS) {   // source row = line, source column = k
                if (l0 + sr < lines && k0 + sc < k) v = X[(l0 + sr) * ld + k0 + sc];
            } else {       // source row = k, source column = line
                if (k0 + sr < k && l0 + sc < lines) v = X[(k0 + sr) * ld + l0 + sc];
            }
            tile[sr * STR + sc] = v;
        }
    }
    __syncthreads();
    const int lane = t & 63;
    const int line = (t >> 6) * 16 + (lane & 15);
    const int kq = lane >> 4;
    const int64_t gl = l0 + line;
    const int e = gl < lines ? ex[gl] : EXP_BAD;
    const int beta = MOD && gl < lines && e != EXP_BAD ? e - lshift[gl] : 0;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const double a = COLS ? tile[(kq * 16 + i) * STR + line] : tile[line * STR + kq * 16 + i];
        const double v = scaled_operand<MOD>(a, e != EXP_BAD ? e : 0, beta);
        x[i] = e != EXP_BAD ? v : 0.0;
    }
    if (BF) {
        // k values 16 kq .. 16 kq + 15 of this 64-k tile: bf16 row block
        // 2 kt + kq / 2, bytes 32 (kq % 2) .. +32
        int8_t *dst = Xop + ((2 * kt + (kq >> 1)) * Lp + gl) * KT + (kq & 1) * 32;
        emit_slices16_bf16(x, planes, dst, KTC * Lp * KT);
        return;
    }
    int8_t *dst = Xop + (kt * Lp + gl) * KT + kq * 16;
    if (MOD) {
        // |X'| <= 2^beta: lines with beta <= 51 take the shorter residue
        // sequence (beta is per line, so a wave may take both)
        if (beta <= 51)
            emit_residues16<true>(x, planes, dst, KTC * Lp * KT);
        else
            emit_residues16<false>(x, planes, dst, KTC * Lp * KT);
    } else
        emit_slices16(x, planes, dst, KTC * Lp * KT);
}

// ---- INT8 product kernel ------------------------------------------------------
//
// One workgroup owns a block of C.  256 x 256: 8 wave64s (2 x 4), two per
// SIMD, each holding a 128 x 64 INT32 accumulator tile; 128 x 128: 4 waves
// (2 x 2) of 64 x 64, two workgroups per CU; 256 x 128 (bf16 form): 8 waves
// (4 x 2) of 64 x 64.  Two waves per SIMD let one wave's LDS reads, barrier
// wait and load issue overlap the other's MFMAs.
//
// Operand staging: a ring of NSTAGE LDS slots, each holding the A and B tiles
// of one 64-byte k stage, is filled by direct global->LDS loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at l*16 of a
// 1 KiB run, no registers involved).  Every step requests exactly one stage
// NSTAGE-1 ahead (NSTAGE-1 stages in flight).  Stages are published with an
// explicit vmcnt wait plus a plain s_barrier (a full __syncthreads would also
// drain the younger stages' loads).  In LDS a tile row is 64 bytes; 16-byte
// chunk c of row r sits at position c ^ swz(r), which keeps the fragment
// reads free of bank conflicts.
//
// Steps run group by group (scheme I: weight group t, slice pairs p + q = t;
// scheme II: one group per modulus), K chunk by K chunk, pair by pair, k tile
// by k tile.  A chunk is a whole even number of k tiles.

constexpr int TGROUP_M = 4;       // C block rows per raster group
constexpr int CHUNK_GRAIN = 2;    // K chunks are whole multiples of this many 64-k tiles
constexpr int NSTAGE = 5;         // LDS ring depth (stages)

// Scheme I group-t accumulation chunk in 64-deep k tiles: the largest whole
// multiple of CHUNK_GRAIN tiles whose INT32 sums over all np pairs of the
// group stay below 2^31.
__device__ inline int chunk_tiles(int np) {
    // floor(floor(a / b) / c) == floor(a / (b c)) for positive integers
    constexpr int PER_PAIR = (int)(2147483647ll / (SLICE_MAX_PRODUCT * KT));
    const int chunk = max(1, PER_PAIR / np);
    return max(1, chunk / CHUNK_GRAIN) * CHUNK_GRAIN;
}

// Position of the next stage to load, walked in the product's step order.
// Scheme I: groups t = 2..s+1, pairs p = max(1,t-s)..min(s,t-1), A plane p-1,
// B plane t-p-1.  Scheme II: groups t = 0..s-1, one pair, A and B plane t.
// The common advance adds one k tile to both operand pointers; group, chunk
// and pair changes take the slow path.  Past the last step the cursor stays
// on the last stage, so the load that keeps the per-step load count uniform
// re-reads a valid tile into a slot no step reads again.
template <bool MOD, bool BF = false>
struct LoadCursor {
    const int8_t *ga, *gb;   // tile base of the current stage (no lane offset)
    int kt, c1;              // current k tile, end of the current chunk
    int t, p, pmax, c0, chunk;
    int s, ktiles;
    int64_t mkt, nkt;        // bytes per k tile of one A / B operand plane
    int64_t a_plane, b_plane;
    const int8_t *abase, *bbase;

    __device__ void place() {
        const int pa = MOD ? t : p - 1;
        const int pb = MOD ? t : t - p - 1;
        ga = abase + (int64_t)pa * a_plane + kt * mkt;
        gb = bbase + (int64_t)pb * b_plane + kt * nkt;
    }
    __device__ int first_pair() const { return MOD ? t : max(1, t - s); }
    __device__ void begin_group() {
        p = first_pair();
        pmax = MOD ? t : min(s, t - 1);
        // bf16 planes hold 32 k values per 64-byte tile: the same k span
        chunk = MOD ? MOD_CHUNK_TILES : (BF ? 2 : 1) * chunk_tiles(pmax - p + 1);
        c0 = 0;
        c1 = min(chunk, ktiles);
        kt = 0;
    }
    __device__ bool last_group() const { return MOD ? t >= s - 1 : t > s; }
    __device__ void start(int s_, int ktiles_, const int8_t *abase_, const int8_t *bbase_,
                          int64_t mkt_, int64_t nkt_, int64_t a_plane_, int64_t b_plane_) {
        s = s_;
        ktiles = ktiles_;
        abase = abase_;
        bbase = bbase_;
        mkt = mkt_;
        nkt = nkt_;
        a_plane = a_plane_;
        b_plane = b_plane_;
        t = MOD ? 0 : 2;
        begin_group();
        place();
    }
    __device__ void advance() {
        if (++kt < c1) {
            ga += mkt;
            gb += nkt;
            return;
        }
        slow_advance();
    }
    __device__ void slow_advance() {
        if (++p <= pmax) {
            kt = c0;
        } else if (c1 < ktiles) {
            c0 = c1;
            c1 = min(c0 + chunk, ktiles);
            p = first_pair();
            kt = c0;
        } else if (!last_group()) {
            ++t;
            begin_group();
        } else {
            kt = c1 - 1;   // past the end: stay on the last stage
            p = pmax;
            return;
        }
        place();
    }
};

// s_waitcnt with vmcnt <= N and the other counters untouched (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// LDS position of the 16-byte chunks of a 64-byte tile row r: chunk c sits at
// c ^ swz(r).  A ds_read_b128 is served 8 lanes (128 bytes) at a time; the 8
// rows of such a group then cover all 32 banks.  Row blocks are 16-row
// aligned, so swz(row) depends on row % 16 only.
__device__ inline int swz(int r) { return (r >> 1) & 3; }

__device__ inline void block_barrier() {
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// 16-byte direct global->LDS load: lane l of the wave writes its bytes at
// lds_run + 16*l (lds_run must be wave-uniform).
__device__ inline void dma16(const int8_t *g, int8_t *lds_run) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void *)lds_run, 16, 0,
                                     0);
}


// Folds one wave's INT32 accumulators into C (scheme I): x = ldexp(G, e_i +
// f_j + shift) (NaN when a row or column exponent is EXP_BAD), C = x on an
// overwriting fold, else C = C + x.  In the INT32 MFMA layout lane l,
// register r of fragment (i, j) holds row 16i + 4(l/16) + r, column
// 16j + l%16 of the wave tile, so one load or store instruction covers four
// 128-byte row segments of C.  row0 / col0 are this lane's first row and
// column.  EDGE adds the bounds checks that tiles crossing the matrix edge
// need.
template <int FI, int FJ, bool EDGE>
__device__ __forceinline__ void fold_tile(const v4i (&acc)[FI][FJ], double *C, int64_t ldc,
                                          const int32_t *rexp, const int32_t *cexp, int64_t M,
                                          int64_t N, int64_t row0, int64_t col0, int shift,
                                          bool overwrite) {
    int ecol[FJ];
#pragma unroll
    for (int j = 0; j < FJ; ++j) ecol[j] = (!EDGE || col0 + 16 * j < N) ? cexp[col0 + 16 * j] : 0;
    double *cptr = C + row0 * ldc + col0;   // row row0 + 16i while fragment row i is folded
    int64_t row = row0;
#pragma unroll
    for (int i = 0; i < FI; ++i) {
        int erow[4];
        double cv[FJ][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool rok = !EDGE || row + r < M;
            erow[r] = rok ? rexp[row + r] : 0;
#pragma unroll
            for (int j = 0; j < FJ; ++j) {
                const bool ok = rok && (!EDGE || col0 + 16 * j < N);
                cv[j][r] = (!overwrite && ok) ? cptr[r * ldc + 16 * j] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int j = 0; j < FJ; ++j) {
                const bool ok = !EDGE || (row + r < M && col0 + 16 * j < N);
                if (ok) {
                    const double x = (erow[r] == EXP_BAD || ecol[j] == EXP_BAD)
                                         ? __longlong_as_double(0x7ff8000000000000ll)
                                         : ldexp((double)acc[i][j][r], erow[r] + ecol[j] + shift);
                    cptr[r * ldc + 16 * j] = overwrite ? x : cv[j][r] + x;
                }
            }
        }
        cptr += 16 * ldc;
        row += 16;
    }
}

// acc mod m in [0, m) for every accumulator (scheme II).  q = floor(a / m) is
// taken from the FP64 quotient, which is within 2^-21 of the exact one, so
// one correction step each way suffices.
template <int FI, int FJ>
__device__ __forceinline__ void reduce_mod(v4i (&acc)[FI][FJ], int m) {
    const double inv = 1.0 / (double)m;
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = acc[i][j][r];
                int v = a - (int)floor((double)a * inv) * m;
                if (v < 0) v += m;
                if (v >= m) v -= m;
                acc[i][j][r] = v;
            }
}

// Layout of the scheme II product residues: per modulus, per C block tile,
// per wave, QPL = FI FJ / 4 runs of 1 KiB; run q holds, for every lane, the
// 16 bytes of fragments 4q..4q+3 (fragment f = i FJ + j, byte 4 (f % 4) + r
// within the lane's 16 bytes is register r).
__device__ __forceinline__ int64_t residue_run(int l, int64_t tiles, int64_t tile, int nw,
                                               int wave, int qpl, int q) {
    return ((((int64_t)l * tiles + tile) * nw + wave) * qpl + q) * 1024;
}

template <int TBM_, int TBN_, int WGM, int WGN>
struct TileShape {
    static constexpr int TBM = TBM_;               // C block rows
    static constexpr int TBN = TBN_;               // C block columns
    static constexpr int NW = WGM * WGN;           // waves per workgroup
    static constexpr int THREADS = 64 * NW;
    static constexpr int WTM = TBM / WGM;          // wave tile rows
    static constexpr int WTN = TBN / WGN;          // wave tile columns
    static constexpr int FI = WTM / 16;            // 16x16 fragments along rows
    static constexpr int FJ = WTN / 16;            // ... along columns
    static constexpr int QPL = FI * FJ / 4;        // 16-byte residue runs per lane
    static constexpr int TILE_A = TBM * KT;        // bytes of the A tile per stage
    static constexpr int TILE_B = TBN * KT;        // bytes of the B tile per stage
    static constexpr int SLOT = TILE_A + TILE_B;   // A tile then B tile
    static constexpr int LPW_A = TILE_A / 1024 / NW;   // 1 KiB loads per wave per stage, A
    static constexpr int LPW_B = TILE_B / 1024 / NW;   // ... B
    static constexpr int LDS_BYTES = NSTAGE * SLOT;
    static_assert(LPW_A >= 1 && LPW_A * NW * 1024 == TILE_A, "loads must cover the A tile");
    static_assert(LPW_B >= 1 && LPW_B * NW * 1024 == TILE_B, "loads must cover the B tile");
    static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");
    static_assert(FI % 2 == 0 && (FI * FJ) % 4 == 0, "fragment grouping");
};
using ShapeLarge = TileShape<256, 256, 2, 4>;
using ShapeSmall = TileShape<128, 128, 2, 2>;
// bf16 form: 8 waves of 64 x 64 (its FP32 and INT32 accumulator sets fit two
// waves per SIMD) in one workgroup per CU.  Two co-resident 128-tile
// workgroups per CU drift apart (the older one's waves win the matrix-core
// issue slots), which spoils the L2 sharing between the workgroups of a
// raster group; one workgroup per CU keeps its waves in step at the barriers.
using ShapeWide = TileShape<256, 128, 4, 2>;

struct ProductArgs {
    const int8_t *A, *B;          // operand planes, k-tile blocked
    const int32_t *rexp, *cexp;   // row / column exponents (scheme I fold)
    double *C;                    // scheme I output
    int64_t ldc;
    uint8_t *R;                   // scheme II product residues
    int64_t M, N, Mp, Np, KTC;
    int s;                        // slices (scheme I) or moduli (scheme II)
    int accumulate;
};

// The product kernels' body.  BF (digit slices only) runs the bf16 planes on
// the bf16 MFMA (16x16x32, FP32 accumulation): same bytes per fragment and
// per LDS stage as INT8, 32 k values per 64-byte tile.  The FP32 sums are
// moved into the INT32 accumulators every BF_FLUSH_STEPS steps, while they
// are still exact integers (16 steps x 32 products x 2^14 = 2^23 < 2^24), so
// the group sums, chunks and folds are those of the INT8 digit-slice form and
// the result is bit-identical to it.
constexpr int BF_FLUSH_STEPS = 16;

template <int TBM, int TBN, int WGM, int WGN, bool MOD, bool BF>
__device__ __forceinline__ void product_body(const ProductArgs &a) {
    static_assert(!(MOD && BF), "bf16 planes carry digit slices only");
    using S = TileShape<TBM, TBN, WGM, WGN>;
    constexpr int FI = S::FI, FJ = S::FJ, TILE_A = S::TILE_A, SLOT = S::SLOT;
    constexpr int LPW_A = S::LPW_A, LPW_B = S::LPW_B;
    // operand ring
    __shared__ __attribute__((aligned(1024))) int8_t lds[S::LDS_BYTES];

    const int64_t M = a.M, N = a.N;
    const int s = a.s;
    const int tiles_m = (int)((M + TBM - 1) / TBM);
    const int tiles_n = (int)((N + TBN - 1) / TBN);
    int tm, tn;
    {
        // XCD-aware remap: consecutive workgroups go to different XCDs, so
        // tile ranges are dealt out per XCD (when the count divides by 8),
        // then rastered in groups of TGROUP_M block rows for L2 reuse.
        const int nb = tiles_m * tiles_n;
        int b = blockIdx.x;
        if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
        const int per_group = TGROUP_M * tiles_n;
        const int g = b / per_group;
        const int first_m = g * TGROUP_M;
        const int gsize = min(tiles_m - first_m, TGROUP_M);
        tm = first_m + (b % per_group) % gsize;
        tn = (b % per_group) / gsize;
    }
    const int64_t m0 = (int64_t)tm * TBM;
    const int64_t n0 = (int64_t)tn * TBN;
    const bool full_tile = m0 + TBM <= M && n0 + TBN <= N;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN;
    const int wn = wave % WGN;
    const int lk = lane >> 4;
    const int lc = lane & 15;

    // Direct-load assignment: load j of this wave fills rows
    // 16*(wave*LPW + j) .. +16 of the A (B) tile, LPW = LPW_A (LPW_B); lane l
    // reads row +l/4, 16-byte chunk (l%4) ^ swz(l/4), landing at position
    // l%4 of that row in LDS.  Consecutive loads are 1 KiB apart in HBM and
    // in LDS.
    const int lane_chunk = (lane >> 2) * KT + (((lane & 3) ^ swz(lane >> 2)) << 4);

    v4i acc[FI][FJ];
    const v4i vzero = {0, 0, 0, 0};
    v4f accf[BF ? FI : 1][BF ? FJ : 1];   // bf16 form: FP32 partial sums
    constexpr int LPS = LPW_A + LPW_B;   // load instructions per wave per stage

    const int ktiles = (int)a.KTC;
    LoadCursor<MOD, BF> cur;
    cur.start(s, ktiles, a.A + m0 * KT, a.B + n0 * KT, a.Mp * KT, a.Np * KT, a.KTC * a.Mp * KT,
              a.KTC * a.Np * KT);
    const uint32_t voff_a = (uint32_t)(16 * wave * LPW_A * KT + lane_chunk);
    const uint32_t voff_b = (uint32_t)(16 * wave * LPW_B * KT + lane_chunk);
    int8_t *const lds_wa = lds + wave * LPW_A * 1024;            // this wave's A load runs
    int8_t *const lds_wb = lds + TILE_A + wave * LPW_B * 1024;   // ... and B load runs
    auto issue = [&](int slot) {
#pragma unroll
        for (int j = 0; j < LPW_A; ++j)
            dma16(cur.ga + voff_a + j * 1024, lds_wa + slot * SLOT + j * 1024);
#pragma unroll
        for (int j = 0; j < LPW_B; ++j)
            dma16(cur.gb + voff_b + j * 1024, lds_wb + slot * SLOT + j * 1024);
        cur.advance();
    };
    // prologue: stages 0 .. NSTAGE-2.  Every step then issues exactly one
    // stage, so NSTAGE-1 stages are in flight at each step's wait.
    for (int st = 0; st < NSTAGE - 1; ++st) issue(st);
    int pslot = NSTAGE - 1;   // ring slot of the most recently published stage

    // Publishes the next stage: this wave's loads of it are done once at most
    // the NSTAGE-2 younger stages are outstanding; after the barrier every
    // wave's loads are in LDS, and every wave has finished reading the
    // previous stage (its fragments fed MFMAs issued before the barrier), so
    // that slot takes the stage NSTAGE-1 ahead.
    auto publish = [&]() {
        const int freed = pslot;
        pslot = pslot + 1 == NSTAGE ? 0 : pslot + 1;
        wait_vmcnt<(NSTAGE - 2) * LPS>();
        block_barrier();
        issue(freed);
    };
    constexpr int FH = FI / 2;   // A fragments per MFMA half
    auto read_lo = [&](v4i (&fa)[FH]) {
        const int8_t *la = lds + pslot * SLOT;
#pragma unroll
        for (int i = 0; i < FH; ++i) {
            const int r = wm * S::WTM + i * 16 + lc;
            fa[i] = *(const v4i *)(la + r * KT + ((lk ^ swz(r)) << 4));
        }
    };
    auto read_hi = [&](v4i (&fa)[FH], v4i (&fb)[FJ]) {
        const int8_t *la = lds + pslot * SLOT;
        const int8_t *lb = la + TILE_A;
#pragma unroll
        for (int i = 0; i < FH; ++i) {
            const int r = wm * S::WTM + (FH + i) * 16 + lc;
            fa[i] = *(const v4i *)(la + r * KT + ((lk ^ swz(r)) << 4));
        }
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
            const int r = wn * S::WTN + j * 16 + lc;
            fb[j] = *(const v4i *)(lb + r * KT + ((lk ^ swz(r)) << 4));
        }
    };
    // Serpentine order (j descending on odd i): consecutive MFMAs share one
    // operand, so only one operand's bits change per instruction.  The part
    // is power-bound under these products; against the row-major order this
    // measured 0.6-1.0% faster (same results: each accumulator sees the same
    // sequence).
    auto mfma_half = [&](int i0, const v4i (&fa)[FH], const v4i (&fb)[FJ]) {
#pragma unroll
        for (int i = 0; i < FH; ++i)
#pragma unroll
            for (int jj = 0; jj < FJ; ++jj) {
                const int j = (i & 1) ? FJ - 1 - jj : jj;
                if constexpr (BF)
                    accf[i0 + i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(v8bf, fa[i]), __builtin_bit_cast(v8bf, fb[j]),
                        accf[i0 + i][j], 0, 0, 0);
                else
                    acc[i0 + i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j],
                                                                           acc[i0 + i][j], 0, 0, 0);
            }
    };
    // Software pipeline: the next step is published and its fragments are
    // read between the two halves of the current step's MFMAs, so barrier
    // and LDS latency overlap matrix work.  The lower A fragments are read
    // into the registers the first half has just consumed; upper A and B
    // fragments alternate between two sets for even and odd steps.
    v4i alo[FH], ahi0[FH], b0[FJ], ahi1[FH], b1[FJ];
    auto run_steps = [&](int nsteps) {   // nsteps even
        publish();
        read_lo(alo);
        read_hi(ahi0, b0);
        for (int n = 0; n < nsteps; n += 2) {
            __builtin_amdgcn_sched_barrier(0);
            mfma_half(0, alo, b0);
            __builtin_amdgcn_sched_barrier(0);
            publish();
            read_lo(alo);
            read_hi(ahi1, b1);
            __builtin_amdgcn_sched_barrier(0);
            mfma_half(FH, ahi0, b0);
            __builtin_amdgcn_sched_barrier(0);
            mfma_half(0, alo, b1);
            __builtin_amdgcn_sched_barrier(0);
            if (n + 2 < nsteps) {
                publish();
                read_lo(alo);
                read_hi(ahi0, b0);
            }
            __builtin_amdgcn_sched_barrier(0);
            mfma_half(FH, ahi1, b1);
        }
    };

    if constexpr (!MOD) {
        bool first_fold = true;
        for (int t = 2; t <= s + 1; ++t) {
            const int np = min(s, t - 1) - max(1, t - s) + 1;
            const int chunk = (BF ? 2 : 1) * chunk_tiles(np);
            for (int c0 = 0; c0 < ktiles; c0 += chunk) {
#pragma unroll
                for (int i = 0; i < FI; ++i)
#pragma unroll
                    for (int j = 0; j < FJ; ++j) acc[i][j] = vzero;
                // chunks and KTC are even, so every chunk has an even step count
                if constexpr (BF) {
                    for (int left = min(chunk, ktiles - c0) * np; left > 0;
                         left -= BF_FLUSH_STEPS) {
#pragma unroll
                        for (int i = 0; i < FI; ++i)
#pragma unroll
                            for (int j = 0; j < FJ; ++j) accf[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
                        run_steps(min(left, BF_FLUSH_STEPS));
#pragma unroll
                        for (int i = 0; i < FI; ++i)
#pragma unroll
                            for (int j = 0; j < FJ; ++j)
#pragma unroll
                                for (int r = 0; r < 4; ++r) acc[i][j][r] += (int)accf[i][j][r];
                    }
                } else {
                    run_steps(min(chunk, ktiles - c0) * np);
                }
                const int64_t row0 = m0 + wm * S::WTM + 4 * lk;
                const int64_t col0 = n0 + wn * S::WTN + lc;
                const bool overwrite = first_fold && !a.accumulate;
                const int shift = 4 - 8 * t;
                if (full_tile)
                    fold_tile<FI, FJ, false>(acc, a.C, a.ldc, a.rexp, a.cexp, M, N, row0, col0,
                                             shift, overwrite);
                else
                    fold_tile<FI, FJ, true>(acc, a.C, a.ldc, a.rexp, a.cexp, M, N, row0, col0,
                                            shift, overwrite);
                first_fold = false;
                // the fold's C traffic retires before the stage counting resumes
                wait_vmcnt<0>();
            }
        }
    } else {
        const int64_t tiles = (int64_t)tiles_m * tiles_n;
        const int64_t tile = (int64_t)tm * tiles_n + tn;
        for (int t = 0; t < s; ++t) {
            const int m = (int)MODULI[t];
#pragma unroll
            for (int i = 0; i < FI; ++i)
#pragma unroll
                for (int j = 0; j < FJ; ++j) acc[i][j] = vzero;
            for (int c0 = 0; c0 < ktiles; c0 += MOD_CHUNK_TILES) {
                run_steps(min(MOD_CHUNK_TILES, ktiles - c0));
                reduce_mod(acc, m);   // keeps the next chunk's INT32 sums in range
            }
#pragma unroll
            for (int q = 0; q < S::QPL; ++q) {
                v4i run;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int f = 4 * q + w;
                    const v4i v = acc[f / FJ][f % FJ];
                    run[w] = (v[0] & 0xff) | ((v[1] & 0xff) << 8) | ((v[2] & 0xff) << 16) |
                             ((v[3] & 0xff) << 24);
                }
                *(v4i *)(a.R + residue_run(t, tiles, tile, S::NW, wave, S::QPL, q) + lane * 16) =
                    run;
            }
            // the residue stores retire before the stage counting resumes
            wait_vmcnt<0>();
        }
    }
}

template <int TB, int WGM, int WGN, bool MOD, int MINB>
__global__ __launch_bounds__(64 * WGM * WGN, MINB) void int8_product_kernel(ProductArgs a) {
    product_body<TB, TB, WGM, WGN, MOD, false>(a);
}

template <int TBM, int TBN, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(64 * WGM * WGN, MINB) void bf16_product_kernel(ProductArgs a) {
    product_body<TBM, TBN, WGM, WGN, false, true>(a);
}

// Scheme II reconstruction: one thread per (block tile, wave, fragment
// quad, lane) of the product kernel's layout, so each wave reads one
// contiguous 1 KiB run per modulus; for each of its 16 outputs it gathers the
// n residues, recovers C' exactly, rounds once and scales.
// NM > 0 fixes the modulus count at compile time (the table words then fold
// into immediates and the unused moduli drop out); NM = 0 reads it from nmod.
// ACC adds the result to C (read before the arithmetic) instead of storing it.
template <int NM>
__device__ __forceinline__ double crt_value(const uint32_t (&res)[MAX_MODULI], int nmod) {
    if constexpr (NM > 0) {
        constexpr mgx::modular::CrtTables T = mgx::modular::make_crt_tables();
        return mgx::modular::reconstruct(T, res, NM);
    } else {
        return mgx::modular::reconstruct(CRT_TABLES, res, nmod);
    }
}

template <int TB, int WGM, int WGN, int NM, bool ACC>
__global__ __launch_bounds__(256) void reconstruct_kernel(const uint8_t *__restrict__ R,
                                                          const int32_t *__restrict__ rshift,
                                                          const int32_t *__restrict__ cshift,
                                                          double *__restrict__ C, int64_t M,
                                                          int64_t N, int64_t ldc, int nmod) {
    using S = TileShape<TB, TB, WGM, WGN>;
    const int nm = NM > 0 ? NM : nmod;
    const int tiles_m = (int)((M + TB - 1) / TB);
    const int tiles_n = (int)((N + TB - 1) / TB);
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= tiles * S::THREADS * S::QPL) return;
    const int lane = (int)(gid % 64);
    const int q = (int)((gid / 64) % S::QPL);
    const int wave = (int)((gid / (64 * S::QPL)) % S::NW);
    const int64_t tile = gid / ((int64_t)S::THREADS * S::QPL);
    const int64_t m0 = (tile / tiles_n) * TB;
    const int64_t n0 = (tile % tiles_n) * TB;
    const int64_t row0 = m0 + (wave / WGN) * S::WTM + 4 * (lane >> 4);
    const int64_t col0 = n0 + (wave % WGN) * S::WTN + (lane & 15);
    // A lane's 16 outputs are fragments f = 4q + w (w = 0..3) at rows
    // 16 (f / FJ) + r and columns 16 (f % FJ): with FJ a multiple of 4 that
    // is 4 rows (r) by 4 columns (w).  Every load (residue runs, line shifts,
    // the accumulated C) is issued before the arithmetic.
    static_assert(S::FJ % 4 == 0, "a residue run covers 4 columns of one fragment row");
    const int64_t rbase = row0 + 16 * ((4 * q) / S::FJ);
    const int64_t cbase = col0 + 16 * ((4 * q) % S::FJ);
    v4i runs[MAX_MODULI];
#pragma unroll
    for (int l = 0; l < MAX_MODULI; ++l)
        if (l < nm)
            runs[l] = *(const v4i *)(R + residue_run(l, tiles, tile, S::NW, wave, S::QPL, q) +
                                     lane * 16);
    int er[4], ec[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) er[r] = rbase + r < M ? rshift[rbase + r] : 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) ec[w] = cbase + 16 * w < N ? cshift[cbase + 16 * w] : 0;
    double cin[4][4];   // the C being accumulated into (ACC only)
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            cin[w][r] = ACC && rbase + r < M && cbase + 16 * w < N
                            ? C[(rbase + r) * ldc + cbase + 16 * w]
                            : 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const int64_t col = cbase + 16 * w;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = rbase + r;
            if (row < M && col < N) {
                uint32_t res[MAX_MODULI];
#pragma unroll
                for (int l = 0; l < MAX_MODULI; ++l)
                    res[l] = l < nm ? ((uint32_t)runs[l][w] >> (8 * r)) & 0xffu : 0u;
                const double x = (er[r] == EXP_BAD || ec[w] == EXP_BAD)
                                     ? __longlong_as_double(0x7ff8000000000000ll)
                                     : ldexp(crt_value<NM>(res, nm), er[r] + ec[w]);
                C[row * ldc + col] = ACC ? cin[w][r] + x : x;
            }
        }
    }
}

__global__ void zero_matrix_kernel(double *C, int64_t m, int64_t n, int64_t ldc) {
    const int64_t total = m * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / n;
        C[r * ldc + (i - r * n)] = 0.0;
    }
}

unsigned grid1d(int64_t total, int threads, int64_t cap = 65536) {
    int64_t b = (total + threads - 1) / threads;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

bool form_planes_ok(int form, int planes) {
    if (form == MGX_FORM_INT8 || form == MGX_FORM_BF16)
        return planes >= MIN_SLICES && planes <= MAX_SLICES;
    if (form == MGX_FORM_INT8_MOD) return planes >= MIN_MODULI && planes <= MAX_MODULI;
    return false;
}

struct SplitLayout {
    int64_t Kp, KTC, Mp, Np;
    size_t a_off, b_off, r_off, rmax_off, cmax_off, rsum_off, csum_off, rexp_off, cexp_off;
    size_t rshift_off, cshift_off, total;
};

SplitLayout split_layout(int form, int planes, int64_t m, int64_t n, int64_t k) {
    SplitLayout L;
    L.Kp = round_up(k < 1 ? 1 : k, K_ALIGN);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    // bytes per plane value: bf16 planes take two
    const int64_t esize = form == MGX_FORM_BF16 ? 2 : 1;
    L.KTC = L.Kp * esize / KT;
    L.Mp = round_up(m < 1 ? 1 : m, ROW_ALIGN);
    L.Np = round_up(n < 1 ? 1 : n, ROW_ALIGN);
    L.a_off = take((size_t)planes * L.Mp * L.Kp * esize);
    L.b_off = take((size_t)planes * L.Np * L.Kp * esize);
    // product residues: every block tile of either tile size fits in Mp x Np
    L.r_off = take(form == MGX_FORM_INT8_MOD ? (size_t)planes * L.Mp * L.Np : 0);
    // row / column maxima, then the modular form's group sums: one block of
    // 2 (m + n) words cleared by one kernel
    L.rmax_off = take((size_t)(m + n) * 16);
    L.cmax_off = L.rmax_off + (size_t)m * 8;
    L.rsum_off = L.cmax_off + (size_t)n * 8;
    L.csum_off = L.rsum_off + (size_t)m * 8;
    L.rexp_off = take((size_t)m * 4);
    L.cexp_off = take((size_t)n * 4);
    L.rshift_off = take((size_t)m * 4);
    L.cshift_off = take((size_t)n * 4);
    L.total = off;
    return L;
}

}  // namespace

extern "C" int mgx_split_slices_range(int form, int *min_slices, int *max_slices) {
    mgx::clear_error();
    int lo, hi;
    if (form == MGX_FORM_INT8 || form == MGX_FORM_BF16) {
        lo = MIN_SLICES;
        hi = MAX_SLICES;
    } else if (form == MGX_FORM_INT8_MOD) {
        lo = MIN_MODULI;
        hi = MAX_MODULI;
    } else {
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "split form not supported");
    }
    if (min_slices) *min_slices = lo;
    if (max_slices) *max_slices = hi;
    return MGX_OK;
}

extern "C" int mgx_split_mod_beta(int moduli, int64_t k) {
    return mgx::modular::beta_for(moduli, k);
}

extern "C" size_t mgx_split_workspace_bytes(int form, int slices, int64_t m, int64_t n,
                                            int64_t k) {
    if (!form_planes_ok(form, slices) || m < 0 || n < 0 || k < 0) return 0;
    return split_layout(form, slices, m, n, k).total;
}

namespace mgx {
int split_gemm_run(int form, int slices, int tile, const double *A, const double *B, double *C,
                   int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                   int accumulate, void *workspace, size_t workspace_bytes, void *stream,
                   void *const *events);
}

static int split_gemm_impl(int form, int slices, int tile, const double *A, const double *B,
                           double *C, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                           int64_t ldc, int accumulate, void *workspace, size_t workspace_bytes,
                           void *stream, void *const *events) {
    mgx::clear_error();
    if (tile == MGX_TILE_DEFAULT) tile = MGX_TILE_LARGE;
    if (tile != MGX_TILE_LARGE && tile != MGX_TILE_SMALL)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "unknown tile");
    if (form != MGX_FORM_INT8 && form != MGX_FORM_INT8_MOD && form != MGX_FORM_BF16)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "split form not supported");
    if (!form_planes_ok(form, slices))
        return mgx::set_error(MGX_ERR_UNSUPPORTED, form == MGX_FORM_INT8_MOD
                                                       ? "modulus count out of range"
                                                       : "slice count out of range");
    const bool mod = form == MGX_FORM_INT8_MOD;
    // the bf16 form runs its own 256 x 128 block tile whatever `tile` asks
    const bool bf = form == MGX_FORM_BF16;
    const int64_t TBm = bf ? ShapeWide::TBM : tile == MGX_TILE_LARGE ? 256 : 128;
    const int64_t TBn = bf ? ShapeWide::TBN : TBm;
    MGX_ARG_CHECK(m >= 0 && n >= 0 && k >= 0, "negative matrix dimension");
    if (m == 0 || n == 0) return MGX_OK;
    MGX_ARG_CHECK(C != nullptr && ldc >= n, "C is NULL or ldc < n");
    hipStream_t st = (hipStream_t)stream;
    if (k == 0) {
        if (accumulate) return MGX_OK;
        hipLaunchKernelGGL(zero_matrix_kernel, dim3(grid1d(m * n, 256)), dim3(256), 0, st, C, m,
                           n, ldc);
        return mgx::launch_status("zero_matrix_kernel");
    }
    MGX_ARG_CHECK(A != nullptr && B != nullptr, "A or B is NULL");
    MGX_ARG_CHECK(lda >= k && ldb >= n, "lda < k or ldb < n");
    // per-line scaling exponents of the modular form: every line gets at
    // least beta_for(slices, k) >= 0
    mgx::modular::BetaThresholds thresholds = {};
    if (mod && (mgx::modular::beta_for(slices, k) < 0 ||
                !mgx::modular::beta_thresholds(slices, k, &thresholds)))
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "too few moduli for this k");
    const SplitLayout L = split_layout(form, slices, m, n, k);
    const int64_t tiles = ((m + TBm - 1) / TBm) * ((n + TBn - 1) / TBn);
    MGX_ARG_CHECK(tiles <= 0x7fffffffll, "matrix too large for one launch");
    // reconstruction: one thread per (tile, wave, fragment quad, lane)
    const int recon_threads = tile == MGX_TILE_LARGE ? ShapeLarge::THREADS * ShapeLarge::QPL
                                                     : ShapeSmall::THREADS * ShapeSmall::QPL;
    const int64_t recon_blocks = (tiles * recon_threads + 255) / 256;
    MGX_ARG_CHECK(!mod || recon_blocks <= 0x7fffffffll, "matrix too large for the reconstruction");
    // operand passes: y = 64-row blocks of A, y = 64-k tiles of B
    MGX_ARG_CHECK(L.Mp / 64 <= 65535 && L.KTC <= 0x7fffffffll, "m or k too large for the row pass");
    MGX_ARG_CHECK(L.Kp / KT <= 65535 && L.Np / 64 <= 0x7fffffffll,
                  "k or n too large for the column pass");
    MGX_ARG_CHECK((k + 255) / 256 <= 65535, "k too large for the column scan");   // seg >= 256 then

    char *ws = (char *)workspace;
    // marks the library scratch as used by `st` once this call's work is queued
    struct ScratchUse {
        bool owned = false;
        void *stream = nullptr;
        ~ScratchUse() {
            if (owned) (void)mgx::scratch_release(mgx::SCRATCH_SPLIT, stream);
        }
    } scratch_use;
    if (ws) {
        if (workspace_bytes < L.total)
            return mgx::set_error(MGX_ERR_WORKSPACE, "split workspace too small");
        if (((uintptr_t)ws & 255) != 0)
            return mgx::set_error(MGX_ERR_INVALID_ARG, "workspace must be 256-byte aligned");
    } else {
        int status = MGX_OK;
        ws = (char *)mgx::device_scratch(mgx::SCRATCH_SPLIT, L.total, stream, &status);
        if (!ws) return status;
        scratch_use.owned = true;
        scratch_use.stream = stream;
    }
    int8_t *Aop = (int8_t *)(ws + L.a_off);
    int8_t *Bop = (int8_t *)(ws + L.b_off);
    uint8_t *Rres = (uint8_t *)(ws + L.r_off);
    unsigned long long *rmax = (unsigned long long *)(ws + L.rmax_off);
    unsigned long long *cmax = (unsigned long long *)(ws + L.cmax_off);
    unsigned long long *rsum = (unsigned long long *)(ws + L.rsum_off);
    unsigned long long *csum = (unsigned long long *)(ws + L.csum_off);
    int32_t *rexp = (int32_t *)(ws + L.rexp_off);
    int32_t *cexp = (int32_t *)(ws + L.cexp_off);
    int32_t *rshift = (int32_t *)(ws + L.rshift_off);
    int32_t *cshift = (int32_t *)(ws + L.cshift_off);
    int rc;
    auto mark = [&](int i) -> int {
        if (events && events[i]) MGX_HIP_TRY(hipEventRecord((hipEvent_t)events[i], st));
        return MGX_OK;
    };

    if ((rc = mark(0)) != MGX_OK) return rc;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1d(2 * (m + n), 256)), dim3(256), 0, st, rmax,
                       2 * (m + n));
    if ((rc = mgx::launch_status("zero_u64_kernel")) != MGX_OK) return rc;
    {
        const unsigned gx = (unsigned)((k + 4095) / 4096);
        const unsigned gy = (unsigned)std::min<int64_t>(m, 65535);
        if (mod)
            hipLaunchKernelGGL(row_scan_kernel<true>, dim3(gx, gy), dim3(256), 0, st, A, m, k, lda,
                               rmax, rsum);
        else
            hipLaunchKernelGGL(row_scan_kernel<false>, dim3(gx, gy), dim3(256), 0, st, A, m, k,
                               lda, rmax, rsum);
        if ((rc = mgx::launch_status("row_scan_kernel")) != MGX_OK) return rc;
    }
    {
        // segments of 256 rows, shorter (down to one 16-row group) when that
        // leaves fewer than 2048 workgroups
        const int64_t gx = (n + 255) / 256;
        int64_t seg = 256;
        while (seg > 16 && gx * ((k + seg - 1) / seg) < 2048) seg /= 2;
        const int64_t gy = (k + seg - 1) / seg;
        if (mod)
            hipLaunchKernelGGL(col_scan_kernel<true>, dim3((unsigned)gx, (unsigned)gy), dim3(256),
                               0, st, B, k, n, ldb, seg, cmax, csum);
        else
            hipLaunchKernelGGL(col_scan_kernel<false>, dim3((unsigned)gx, (unsigned)gy), dim3(256),
                               0, st, B, k, n, ldb, seg, cmax, csum);
        if ((rc = mgx::launch_status("col_scan_kernel")) != MGX_OK) return rc;
    }
    if (mod) {
        hipLaunchKernelGGL(line_scales_kernel<true>, dim3(grid1d(m, 256)), dim3(256), 0, st, rmax,
                           rsum, rexp, rshift, m, thresholds);
        hipLaunchKernelGGL(line_scales_kernel<true>, dim3(grid1d(n, 256)), dim3(256), 0, st, cmax,
                           csum, cexp, cshift, n, thresholds);
    } else {
        hipLaunchKernelGGL(line_scales_kernel<false>, dim3(grid1d(m, 256)), dim3(256), 0, st, rmax,
                           rsum, rexp, rshift, m, thresholds);
        hipLaunchKernelGGL(line_scales_kernel<false>, dim3(grid1d(n, 256)), dim3(256), 0, st, cmax,
                           csum, cexp, cshift, n, thresholds);
    }
    if ((rc = mgx::launch_status("line_scales_kernel")) != MGX_OK) return rc;
    if ((rc = mark(1)) != MGX_OK) return rc;

    {
        const int64_t gy = L.Mp / 64;
        // one workgroup per 64 k values and 64 lines
        const dim3 grid((unsigned)(L.Kp / KT), (unsigned)gy);
        if (bf)
            hipLaunchKernelGGL((operand_tile_kernel<false, false, true>), grid, dim3(256), 0, st,
                               A, m, k, lda, rexp, rshift, slices, L.Mp, L.KTC, Aop);
        else if (mod)
            hipLaunchKernelGGL((operand_tile_kernel<true, false>), grid, dim3(256), 0, st, A, m, k,
                               lda, rexp, rshift, slices, L.Mp, L.KTC, Aop);
        else
            hipLaunchKernelGGL((operand_tile_kernel<false, false>), grid, dim3(256), 0, st, A, m, k,
                               lda, rexp, rshift, slices, L.Mp, L.KTC, Aop);
        if ((rc = mgx::launch_status("operand_tile_kernel (A)")) != MGX_OK) return rc;
    }
    {
        // x: 64-column blocks of B, y: k tiles
        const dim3 grid((unsigned)(L.Np / 64), (unsigned)(L.Kp / KT));
        if (bf)
            hipLaunchKernelGGL((operand_tile_kernel<false, true, true>), grid, dim3(256), 0, st,
                               B, n, k, ldb, cexp, cshift, slices, L.Np, L.KTC, Bop);
        else if (mod)
            hipLaunchKernelGGL((operand_tile_kernel<true, true>), grid, dim3(256), 0, st, B, n, k,
                               ldb, cexp, cshift, slices, L.Np, L.KTC, Bop);
        else
            hipLaunchKernelGGL((operand_tile_kernel<false, true>), grid, dim3(256), 0, st, B, n, k,
                               ldb, cexp, cshift, slices, L.Np, L.KTC, Bop);
        if ((rc = mgx::launch_status("operand_tile_kernel (B)")) != MGX_OK) return rc;
    }
    if ((rc = mark(2)) != MGX_OK) return rc;
    ProductArgs pa;
    pa.A = Aop;
    pa.B = Bop;
    pa.rexp = rexp;
    pa.cexp = cexp;
    pa.C = C;
    pa.ldc = ldc;
    pa.R = Rres;
    pa.M = m;
    pa.N = n;
    pa.Mp = L.Mp;
    pa.Np = L.Np;
    pa.KTC = L.KTC;
    pa.s = slices;
    pa.accumulate = accumulate ? 1 : 0;
    if (bf) {
        hipLaunchKernelGGL((bf16_product_kernel<256, 128, 4, 2, 1>), dim3((unsigned)tiles),
                           dim3(ShapeWide::THREADS), 0, st, pa);
    } else if (tile == MGX_TILE_LARGE) {
        if (mod)
            hipLaunchKernelGGL((int8_product_kernel<256, 2, 4, true, 1>), dim3((unsigned)tiles),
                               dim3(ShapeLarge::THREADS), 0, st, pa);
        else
            hipLaunchKernelGGL((int8_product_kernel<256, 2, 4, false, 1>), dim3((unsigned)tiles),
                               dim3(ShapeLarge::THREADS), 0, st, pa);
    } else {
        if (mod)
            hipLaunchKernelGGL((int8_product_kernel<128, 2, 2, true, 2>), dim3((unsigned)tiles),
                               dim3(ShapeSmall::THREADS), 0, st, pa);
        else
            hipLaunchKernelGGL((int8_product_kernel<128, 2, 2, false, 2>), dim3((unsigned)tiles),
                               dim3(ShapeSmall::THREADS), 0, st, pa);
    }
    if ((rc = mgx::launch_status("int8_product_kernel")) != MGX_OK) return rc;
    if ((rc = mark(3)) != MGX_OK) return rc;
    if (mod) {
        const int64_t blocks = recon_blocks;
        // the common modulus counts get compile-time specialisations
        auto launch = [&](auto tb, auto nm) {
            constexpr int TBv = decltype(tb)::value, NMv = decltype(nm)::value;
            constexpr int WGNv = TBv == 256 ? 4 : 2;
            if (accumulate)
                hipLaunchKernelGGL((reconstruct_kernel<TBv, 2, WGNv, NMv, true>),
                                   dim3((unsigned)blocks), dim3(256), 0, st, Rres, rshift, cshift,
                                   C, m, n, ldc, slices);
            else
                hipLaunchKernelGGL((reconstruct_kernel<TBv, 2, WGNv, NMv, false>),
                                   dim3((unsigned)blocks), dim3(256), 0, st, Rres, rshift, cshift,
                                   C, m, n, ldc, slices);
        };
        using L = std::integral_constant<int, 256>;
        using Sm = std::integral_constant<int, 128>;
        auto by_count = [&](auto tb) {
            if (slices == 15) launch(tb, std::integral_constant<int, 15>{});
            else if (slices == 16) launch(tb, std::integral_constant<int, 16>{});
            else if (slices == 14) launch(tb, std::integral_constant<int, 14>{});
            else launch(tb, std::integral_constant<int, 0>{});
        };
        if (tile == MGX_TILE_LARGE) by_count(L{});
        else by_count(Sm{});
        if ((rc = mgx::launch_status("reconstruct_kernel")) != MGX_OK) return rc;
    }
    return MGX_OK;
}

extern "C" int mgx_split_gemm(int form, int slices, const double *A, const double *B, double *C,
                              int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                              int64_t ldc, int accumulate, void *workspace,
                              size_t workspace_bytes, void *stream) {
    return split_gemm_impl(form, slices, MGX_TILE_DEFAULT, A, B, C, m, n, k, lda, ldb, ldc,
                           accumulate, workspace, workspace_bytes, stream, nullptr);
}

extern "C" int mgx_split_gemm_events(int form, int slices, const double *A, const double *B,
                                     double *C, int64_t m, int64_t n, int64_t k, int64_t lda,
                                     int64_t ldb, int64_t ldc, int accumulate, void *workspace,
                                     size_t workspace_bytes, void *stream, void *const *events) {
    return split_gemm_impl(form, slices, MGX_TILE_DEFAULT, A, B, C, m, n, k, lda, ldb, ldc,
                           accumulate, workspace, workspace_bytes, stream, events);
}

extern "C" int mgx_split_gemm_ex(int form, int slices, int tile, const double *A, const double *B,
                                 double *C, int64_t m, int64_t n, int64_t k, int64_t lda,
                                 int64_t ldb, int64_t ldc, int accumulate, void *workspace,
                                 size_t workspace_bytes, void *stream) {
    return split_gemm_impl(form, slices, tile, A, B, C, m, n, k, lda, ldb, ldc, accumulate,
                           workspace, workspace_bytes, stream, nullptr);
}

int mgx::split_gemm_run(int form, int slices, int tile, const double *A, const double *B,
                        double *C, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                        int64_t ldc, int accumulate, void *workspace, size_t workspace_bytes,
                        void *stream, void *const *events) {
    return split_gemm_impl(form, slices, tile, A, B, C, m, n, k, lda, ldb, ldc, accumulate,
                           workspace, workspace_bytes, stream, events);
}
