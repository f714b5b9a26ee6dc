// FP64-accurate split-precision GEMM (Ozaki scheme I) on the INT8 matrix
// cores of gfx950.
//
// No counterpart in the reference: it computes the same local product as
// ABMultiply / ABMultiplyAccumulateAsync (1D-row-block/mulmat_kernel.cu:86,
// 148) through a different arithmetic.
//
// Algorithm (INT8 form):
//  1. e_i = frexp exponent of max_k |A_ik| (so |A_ik| < 2^e_i), f_j likewise
//     for column j of B.  A row/column holding Inf/NaN gets a sentinel and
//     its outputs become NaN.
//  2. x = A_ik * 2^-e_i in (-1,1) is cut into s signed digits by round to
//     nearest: sigma_1 = rint(x 2^6), then sigma_p = rint(256 r) of the
//     previous remainder r, with a carry pass that keeps every digit in
//     [-128, 127] (slice_digits).  Place values are 2^(2-8p); the residual
//     after s slices is at most 2^(1-8s).  Every step is exact in FP64.  B is
//     sliced per column with f_j and stored transposed (n-major, k contiguous).
//  3. For each weight group t = 2..s+1 (slice pairs p+q = t, place value
//     2^(4-8t)), the INT8 MFMA accumulates G_t = sum_{p+q=t} sum_k sigmaA_p
//     sigmaB_q exactly in INT32.  The K range is cut into chunks of `chunk`
//     64-deep tiles (whole LDS stages), chosen so that
//     pairs(t) * chunk * 64 * 128^2 < 2^31.
//  4. After each chunk the sums are folded into C in FP64:
//     x = ldexp((double)G, e_i + f_j + 4 - 8t) (exact); the first fold writes
//     C = x (or C_in + x for the accumulate form), later folds C = C + x.
// Pairs with p+q > s+1 are dropped; with the residuals they bound the error
// by about (s+2) 2^(2-8s) K 2^(e_i+f_j).
#include "mgx_internal.hpp"

#include <algorithm>

namespace {

constexpr int KT = 64;             // k depth of one INT8 MFMA step / LDS stage
constexpr int K_ALIGN = 256;       // slice rows are zero-padded to this multiple
constexpr int EXP_BAD = 0x40000000;
constexpr int MIN_SLICES = 1;
constexpr int MAX_SLICES = 9;
constexpr long long SLICE_MAX_PRODUCT = 128ll * 128ll;  // largest |sigma_p sigma_q|

typedef int v4i __attribute__((ext_vector_type(4)));

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// ---- exponent scan --------------------------------------------------------

__global__ void zero_u64_kernel(unsigned long long *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__device__ inline unsigned long long abs_bits(double x) {
    return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull;
}

// Row maxima of A (m x k, lda): grid.x covers k in segments of 256*16,
// grid.y strides over rows.  Maxima are kept as the bit patterns of |x|,
// which order like the values for non-negative doubles (NaN above Inf).
__global__ __launch_bounds__(256) void row_absmax_kernel(const double *__restrict__ A, int64_t m,
                                                         int64_t k, int64_t lda,
                                                         unsigned long long *rowmax) {
    __shared__ unsigned long long sh[4];
    const int64_t kbase = (int64_t)blockIdx.x * (256 * 16);
    for (int64_t r = blockIdx.y; r < m; r += gridDim.y) {
        unsigned long long mx = 0;
        const double *row = A + r * lda;
#pragma unroll 4
        for (int i = 0; i < 16; ++i) {
            const int64_t kk = kbase + i * 256 + threadIdx.x;
            if (kk < k) mx = max(mx, abs_bits(row[kk]));
        }
        for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned long long)__shfl_xor(mx, off));
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            mx = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
            if (mx) atomicMax(rowmax + r, mx);
        }
        __syncthreads();
    }
}

// Column maxima of B (k x n, ldb): one lane per column, grid.y splits k in
// segments of 256 rows.
__global__ __launch_bounds__(256) void col_absmax_kernel(const double *__restrict__ B, int64_t k,
                                                         int64_t n, int64_t ldb,
                                                         unsigned long long *colmax) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int64_t k0 = (int64_t)blockIdx.y * 256;
    const int64_t k1 = min(k, k0 + 256);
    unsigned long long mx = 0;
    for (int64_t kk = k0; kk < k1; ++kk) mx = max(mx, abs_bits(B[kk * ldb + c]));
    if (mx) atomicMax(colmax + c, mx);
}

__device__ inline int exponent_of_bits(unsigned long long bits) {
    if (bits == 0) return 0;
    if (bits >= 0x7ff0000000000000ull) return EXP_BAD;
    int e;
    frexp(__longlong_as_double((long long)bits), &e);
    return e;
}

__global__ void exponents_kernel(const unsigned long long *maxbits, int32_t *e, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        e[i] = exponent_of_bits(maxbits[i]);
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

// ---- slice layout ------------------------------------------------------------
//
// Slices are stored k-tile blocked: slice p of A is KTC = Kp/64 blocks, block kt
// holding rows 0..Mp-1 (Mp = M rounded up to 256, padded rows zero) x 64 bytes
// of k:   Asl[((p * KTC + kt) * Mp + row) * 64 + (k % 64)].
// B is stored the same way per column of B (n-major, k contiguous).  A 16-row x
// 64-byte piece of a block is therefore one contiguous KiB in HBM.
constexpr int ROW_ALIGN = 256;

// One workgroup slices 64 rows x 64 k of A: lane t takes row t/4 and the 16 k
// values (t%4)*16.. of it, so 64 lanes write one contiguous KiB per slice.
__global__ __launch_bounds__(256) void slice_rows_kernel(const double *__restrict__ A, int64_t m,
                                                         int64_t k, int64_t lda,
                                                         const int32_t *__restrict__ rexp, int s,
                                                         int64_t Mp, int64_t KTC,
                                                         int8_t *__restrict__ Asl) {
    const int64_t kt = blockIdx.x;
    const int64_t row = (int64_t)blockIdx.y * 64 + (threadIdx.x >> 2);
    const int kq = threadIdx.x & 3;
    const int64_t k0 = kt * KT + kq * 16;
    const int e = row < m ? rexp[row] : EXP_BAD;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t kk = k0 + i;
        x[i] = (row < m && kk < k && e != EXP_BAD) ? ldexp(A[row * lda + kk], -e) : 0.0;
    }
    emit_slices16(x, s, Asl + (kt * Mp + row) * KT + kq * 16, KTC * Mp * KT);
}

// One workgroup slices 64 columns x 64 k of B: lane t takes column t/4 and the
// 16 k values (t%4)*16.. of it.
__global__ __launch_bounds__(256) void slice_cols_kernel(const double *__restrict__ B, int64_t k,
                                                         int64_t n, int64_t ldb,
                                                         const int32_t *__restrict__ cexp, int s,
                                                         int64_t Np, int64_t KTC,
                                                         int8_t *__restrict__ Bsl) {
    const int64_t kt = blockIdx.x;
    const int64_t c = (int64_t)blockIdx.y * 64 + (threadIdx.x >> 2);
    const int kq = threadIdx.x & 3;
    const int64_t k0 = kt * KT + kq * 16;
    const int e = c < n ? cexp[c] : EXP_BAD;
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t kk = k0 + i;
        x[i] = (c < n && kk < k && e != EXP_BAD) ? ldexp(B[kk * ldb + c], -e) : 0.0;
    }
    emit_slices16(x, s, Bsl + (kt * Np + c) * KT + kq * 16, KTC * Np * KT);
}

// ---- fused slice-pair GEMM --------------------------------------------------
//
// One workgroup of 256 lanes (4 wave64s, 2x2) owns a TB x TB block of C (TB =
// 256 or 128); each wave owns (TB/2)^2 as 16x16 INT32 MFMA accumulators (256
// accumulation registers per lane at TB = 256).
//
// Operand staging: a ring of NSTAGE LDS slots, each holding the A and B tiles
// of one 64-byte k stage, is filled by direct global->LDS loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at l*16 of a
// 1 KiB run, no registers involved).  Stage s+NSTAGE-1 is requested while
// stage s is multiplied, so NSTAGE-1 stages of loads are in flight.  Stages are
// published with an explicit vmcnt wait plus a plain s_barrier (a full
// __syncthreads would also drain the younger stages' loads).
// In LDS a tile row is 64 bytes; 16-byte chunk c of row r sits at position
// c ^ ((r >> 2) & 3), so the 16-byte fragment reads of 16 consecutive rows hit
// 16 distinct bank groups.
//
// Steps run in (group t, K chunk, pair p, k stage) order.  At the end of every
// K chunk the INT32 sums are folded into C in FP64:
//     x = ldexp((double)G, e_i + f_j + 4 - 8t)      (exact)
//     C = x  or  C_in + x   on the first fold (plain / accumulate form),
//     C = C + x             on every later fold.
// so C is read and written once per fold; C stays in HBM between folds.

constexpr int TNT = 256;          // lanes per workgroup
constexpr int TGROUP_M = 8;       // C block rows per raster group
constexpr int NSTAGE = 4;         // LDS ring depth (stages)
constexpr int CHUNK_GRAIN = 2;    // K chunks are whole multiples of this many 64-k tiles

// Walks the (group t, K chunk, pair p, k tile) steps in execution order.
struct StepIter {
    int s, ktiles;
    int t, pmin, pmax, p, c0, c1, kt, chunk;

    __device__ void start(int s_, int ktiles_) {
        s = s_;
        ktiles = ktiles_;
        t = 2;
        begin_group();
    }
    __device__ void begin_group() {
        pmin = max(1, t - s);
        pmax = min(s, t - 1);
        const long long np = pmax - pmin + 1;
        chunk = (int)max(1ll, 2147483647ll / (np * SLICE_MAX_PRODUCT * KT));
        chunk = max(1, chunk / CHUNK_GRAIN) * CHUNK_GRAIN;
        c0 = 0;
        c1 = min(chunk, ktiles);
        p = pmin;
        kt = 0;
    }
    // Moves past the current step.  Returns true when that step closed an
    // accumulation chunk of the group that was current before the call.
    __device__ bool advance() {
        if (++kt < c1) return false;
        if (++p <= pmax) {
            kt = c0;
            return false;
        }
        c0 = c1;
        if (c0 < ktiles) {
            c1 = min(c0 + chunk, ktiles);
            p = pmin;
            kt = c0;
            return true;
        }
        ++t;
        if (t <= s + 1) begin_group();
        return true;
    }
};

// s_waitcnt with vmcnt <= N and the other counters untouched (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

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

template <int TB>
__global__ __launch_bounds__(TNT, TB == 256 ? 1 : 2) void ozaki_int8_gemm(
    const int8_t *__restrict__ Asl, const int8_t *__restrict__ Bsl,
    const int32_t *__restrict__ rexp, const int32_t *__restrict__ cexp, double *__restrict__ C,
    int64_t M, int64_t N, int64_t Mp, int64_t Np, int64_t KTC, int64_t ldc, int s, int64_t nsteps,
    int accumulate) {
    constexpr int WT = TB / 2;                 // wave tile edge
    constexpr int FI = WT / 16;                // 16x16 fragments per wave edge
    constexpr int TILE = TB * KT;              // bytes of one operand tile per stage
    constexpr int SLOT = 2 * TILE;             // A tile then B tile
    constexpr int LPW = TILE / 1024 / 4;       // 1 KiB loads per wave per operand per stage
    constexpr int FOLD_STR = WT + 4;           // ints per row of a wave's fold scratch
    constexpr int FOLD_ROWS = 8;               // C rows per fold pass
    static_assert(4 * FOLD_ROWS * FOLD_STR * 4 <= SLOT, "fold scratch must fit one LDS slot");
    __shared__ __attribute__((aligned(1024))) int8_t lds[NSTAGE * SLOT];

    const int tiles_m = (int)((M + TB - 1) / TB);
    const int tiles_n = (int)((N + TB - 1) / TB);
    int tm, tn;
    {
        const int nb = tiles_m * tiles_n;
        int b = blockIdx.x;
        if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
        const int per_group = TGROUP_M * tiles_n;
        const int g = b / per_group;
        const int first_m = g * TGROUP_M;
        const int gsize = min(tiles_m - first_m, TGROUP_M);
        const int r = b - g * per_group;
        tm = first_m + r % gsize;
        tn = r / gsize;
    }
    const int64_t m0 = (int64_t)tm * TB;
    const int64_t n0 = (int64_t)tn * TB;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1;
    const int wn = wave & 1;
    const int lk = lane >> 4;
    const int lc = lane & 15;

    // Direct-load assignment: load j of this wave fills tile rows
    // 16*(wave*LPW + j) .. +16; lane l reads row +l/4, 16-byte chunk
    // (l%4) ^ ((row>>2)&3) = (l%4) ^ ((l>>4)&3), landing at position l%4 of
    // that row in LDS.  Consecutive loads are 1 KiB apart in HBM and in LDS.
    const int lane_off = (16 * wave * LPW + (lane >> 2)) * KT + (((lane & 3) ^ ((lane >> 4) & 3)) << 4);
    const int64_t a_slice = KTC * Mp * KT;
    const int64_t b_slice = KTC * Np * KT;
    auto issue = [&](int slot, int p, int q, int kt) {
        const int8_t *ga = Asl + (int64_t)(p - 1) * a_slice + ((int64_t)kt * Mp + m0) * KT + lane_off;
        const int8_t *gb = Bsl + (int64_t)(q - 1) * b_slice + ((int64_t)kt * Np + n0) * KT + lane_off;
        int8_t *la = lds + slot * SLOT + wave * LPW * 1024;
        int8_t *lb = la + TILE;
#pragma unroll
        for (int j = 0; j < LPW; ++j) {
            dma16(ga + j * 1024, la + j * 1024);
            dma16(gb + j * 1024, lb + j * 1024);
        }
    };

    v4i acc[FI][FI];
    const v4i vzero = {0, 0, 0, 0};

    const int ktiles = (int)KTC;
    StepIter comp, load;
    comp.start(s, ktiles);
    load.start(s, ktiles);
    // prologue: stages 0 .. NSTAGE-2
    for (int st = 0; st < NSTAGE - 1 && st < nsteps; ++st) {
        issue(st, load.p, load.t - load.p, load.kt);
        load.advance();
    }
    bool first_fold = true;
    constexpr int LPS = 2 * LPW;   // load instructions per wave per stage

    int64_t step = 0;
    while (step < nsteps) {
        // one accumulation chunk: steps until the iterator closes the chunk
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
            for (int j = 0; j < FI; ++j) acc[i][j] = vzero;
        int t_cur;
        bool fold;
        do {
            const int slot = (int)(step % NSTAGE);
            // this wave's loads of `step` are done once at most the younger
            // stages' loads are outstanding
            const int64_t younger = min((int64_t)(NSTAGE - 2), nsteps - 1 - step);
            if (younger >= 2)
                wait_vmcnt<2 * LPS>();
            else if (younger == 1)
                wait_vmcnt<LPS>();
            else
                wait_vmcnt<0>();
            block_barrier();   // every wave's loads of `step` are in LDS; slot step-1 is free
            if (step + NSTAGE - 1 < nsteps) {
                issue((int)((step + NSTAGE - 1) % NSTAGE), load.p, load.t - load.p, load.kt);
                load.advance();
            }
            t_cur = comp.t;
            fold = comp.advance();

            const int8_t *la = lds + slot * SLOT;
            const int8_t *lb = la + TILE;
            v4i af[FI], bf[FI];
#pragma unroll
            for (int i = 0; i < FI; ++i) {
                const int r = wm * WT + i * 16 + lc;
                af[i] = *(const v4i *)(la + r * KT + ((lk ^ ((r >> 2) & 3)) << 4));
            }
#pragma unroll
            for (int j = 0; j < FI; ++j) {
                const int r = wn * WT + j * 16 + lc;
                bf[j] = *(const v4i *)(lb + r * KT + ((lk ^ ((r >> 2) & 3)) << 4));
            }
#pragma unroll
            for (int i = 0; i < FI; ++i)
#pragma unroll
                for (int j = 0; j < FI; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
            ++step;
        } while (!fold);

        // Fold through LDS slot (step-1) % NSTAGE: its operands are consumed
        // once every wave passes the barrier; the in-flight loads target other
        // slots.  Pass (i, h) dumps fragment row i, rows 8h..8h+7 (lanes with
        // lk/2 == h), then folds them into C, WT/64 adjacent columns per lane.
        {
            constexpr int CPL = WT / 64;
            __syncthreads();
            const int slot = (int)((step - 1) % NSTAGE);
            int *fold_buf = (int *)(lds + slot * SLOT) + wave * (FOLD_ROWS * FOLD_STR);
            int64_t fcol[CPL];
            int ecol[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                fcol[c] = n0 + wn * WT + CPL * lane + c;
                ecol[c] = fcol[c] < N ? cexp[fcol[c]] : 0;
            }
#pragma unroll
            for (int i = 0; i < FI; ++i) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    // INT32 layout: lane l, register r holds row 4*(l/16)+r, column l%16.
                    if ((lk >> 1) == h) {
#pragma unroll
                        for (int j = 0; j < FI; ++j)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                fold_buf[((lk & 1) * 4 + r) * FOLD_STR + j * 16 + lc] = acc[i][j][r];
                    }
                    __syncthreads();
#pragma unroll 2
                    for (int rho = 0; rho < FOLD_ROWS; ++rho) {
                        const int64_t row = m0 + wm * WT + i * 16 + h * 8 + rho;
                        if (row >= M) break;
                        const int er = rexp[row];
                        double *c_row = C + row * ldc;
#pragma unroll
                        for (int c = 0; c < CPL; ++c) {
                            if (fcol[c] < N) {
                                const int g = fold_buf[rho * FOLD_STR + CPL * lane + c];
                                const double x = (er == EXP_BAD || ecol[c] == EXP_BAD)
                                                     ? __longlong_as_double(0x7ff8000000000000ll)
                                                     : ldexp((double)g, er + ecol[c] + 4 - 8 * t_cur);
                                c_row[fcol[c]] = (first_fold && !accumulate) ? x : c_row[fcol[c]] + x;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            first_fold = false;
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

struct SplitLayout {
    int64_t Kp, KTC, Mp, Np;
    size_t a_off, b_off, rmax_off, cmax_off, rexp_off, cexp_off, total;
};

SplitLayout split_layout(int s, int64_t m, int64_t n, int64_t k) {
    SplitLayout L;
    L.Kp = round_up(k < 1 ? 1 : k, K_ALIGN);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    L.KTC = L.Kp / KT;
    L.Mp = round_up(m < 1 ? 1 : m, ROW_ALIGN);
    L.Np = round_up(n < 1 ? 1 : n, ROW_ALIGN);
    L.a_off = take((size_t)s * L.Mp * L.Kp);
    L.b_off = take((size_t)s * L.Np * L.Kp);
    L.rmax_off = take((size_t)(m + n) * 8);
    L.cmax_off = L.rmax_off + (size_t)m * 8;
    L.rexp_off = take((size_t)m * 4);
    L.cexp_off = take((size_t)n * 4);
    L.total = off;
    return L;
}

}  // namespace

extern "C" int mgx_split_slices_range(int form, int *min_slices, int *max_slices) {
    mgx::clear_error();
    if (form != MGX_FORM_INT8)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "split form not supported");
    if (min_slices) *min_slices = MIN_SLICES;
    if (max_slices) *max_slices = MAX_SLICES;
    return MGX_OK;
}

extern "C" size_t mgx_split_workspace_bytes(int form, int slices, int64_t m, int64_t n,
                                            int64_t k) {
    if (form != MGX_FORM_INT8 || slices < MIN_SLICES || slices > MAX_SLICES || m < 0 || n < 0 ||
        k < 0)
        return 0;
    return split_layout(slices, m, n, k).total;
}

static int split_gemm_impl(int form, int slices, int tile, const double *A, const double *B,
                           double *C, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                           int64_t ldc, int accumulate, void *workspace, size_t workspace_bytes,
                           void *stream, void *const *events) {
    mgx::clear_error();
    if (tile == MGX_TILE_DEFAULT) tile = MGX_TILE_LARGE;
    if (tile != MGX_TILE_LARGE && tile != MGX_TILE_SMALL)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "unknown tile");
    const int64_t TBsz = tile == MGX_TILE_LARGE ? 256 : 128;
    if (form != MGX_FORM_INT8)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "split form not supported");
    if (slices < MIN_SLICES || slices > MAX_SLICES)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "slice count out of range");
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
    const SplitLayout L = split_layout(slices, m, n, k);
    const int64_t tiles = ((m + TBsz - 1) / TBsz) * ((n + TBsz - 1) / TBsz);
    MGX_ARG_CHECK(tiles <= 0x7fffffffll, "matrix too large for one launch");

    char *ws = (char *)workspace;
    if (ws) {
        if (workspace_bytes < L.total)
            return mgx::set_error(MGX_ERR_WORKSPACE, "split workspace too small");
        if (((uintptr_t)ws & 255) != 0)
            return mgx::set_error(MGX_ERR_INVALID_ARG, "workspace must be 256-byte aligned");
    } else {
        int status = MGX_OK;
        ws = (char *)mgx::device_scratch(mgx::SCRATCH_SPLIT, L.total, &status);
        if (!ws) return status;
    }
    int8_t *Asl = (int8_t *)(ws + L.a_off);
    int8_t *Bsl = (int8_t *)(ws + L.b_off);
    unsigned long long *rmax = (unsigned long long *)(ws + L.rmax_off);
    unsigned long long *cmax = (unsigned long long *)(ws + L.cmax_off);
    int32_t *rexp = (int32_t *)(ws + L.rexp_off);
    int32_t *cexp = (int32_t *)(ws + L.cexp_off);
    int rc;
    auto mark = [&](int i) -> int {
        if (events && events[i]) MGX_HIP_TRY(hipEventRecord((hipEvent_t)events[i], st));
        return MGX_OK;
    };

    if ((rc = mark(0)) != MGX_OK) return rc;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(grid1d(m + n, 256)), dim3(256), 0, st, rmax, m + n);
    if ((rc = mgx::launch_status("zero_u64_kernel")) != MGX_OK) return rc;
    {
        const unsigned gx = (unsigned)((k + 4095) / 4096);
        const unsigned gy = (unsigned)std::min<int64_t>(m, 65535);
        hipLaunchKernelGGL(row_absmax_kernel, dim3(gx, gy), dim3(256), 0, st, A, m, k, lda, rmax);
        if ((rc = mgx::launch_status("row_absmax_kernel")) != MGX_OK) return rc;
    }
    {
        const unsigned gx = (unsigned)((n + 255) / 256);
        const int64_t gy = (k + 255) / 256;
        MGX_ARG_CHECK(gy <= 65535, "k too large for the column scan");
        hipLaunchKernelGGL(col_absmax_kernel, dim3(gx, (unsigned)gy), dim3(256), 0, st, B, k, n,
                           ldb, cmax);
        if ((rc = mgx::launch_status("col_absmax_kernel")) != MGX_OK) return rc;
    }
    hipLaunchKernelGGL(exponents_kernel, dim3(grid1d(m, 256)), dim3(256), 0, st, rmax, rexp, m);
    if ((rc = mgx::launch_status("exponents_kernel")) != MGX_OK) return rc;
    hipLaunchKernelGGL(exponents_kernel, dim3(grid1d(n, 256)), dim3(256), 0, st, cmax, cexp, n);
    if ((rc = mgx::launch_status("exponents_kernel")) != MGX_OK) return rc;
    if ((rc = mark(1)) != MGX_OK) return rc;

    {
        const int64_t gy = L.Mp / 64;
        MGX_ARG_CHECK(gy <= 65535 && L.KTC <= 0x7fffffffll, "m or k too large for the row slicer");
        hipLaunchKernelGGL(slice_rows_kernel, dim3((unsigned)L.KTC, (unsigned)gy), dim3(256), 0, st,
                           A, m, k, lda, rexp, slices, L.Mp, L.KTC, Asl);
        if ((rc = mgx::launch_status("slice_rows_kernel")) != MGX_OK) return rc;
    }
    {
        const int64_t gy = L.Np / 64;
        MGX_ARG_CHECK(gy <= 65535, "n too large for the column slicer");
        hipLaunchKernelGGL(slice_cols_kernel, dim3((unsigned)L.KTC, (unsigned)gy), dim3(256), 0, st,
                           B, k, n, ldb, cexp, slices, L.Np, L.KTC, Bsl);
        if ((rc = mgx::launch_status("slice_cols_kernel")) != MGX_OK) return rc;
    }
    if ((rc = mark(2)) != MGX_OK) return rc;
    const int64_t nsteps = L.KTC * (int64_t)(slices * (slices + 1) / 2);
    if (tile == MGX_TILE_LARGE)
        hipLaunchKernelGGL(ozaki_int8_gemm<256>, dim3((unsigned)tiles), dim3(TNT), 0, st, Asl, Bsl,
                           rexp, cexp, C, m, n, L.Mp, L.Np, L.KTC, ldc, slices, nsteps,
                           accumulate ? 1 : 0);
    else
        hipLaunchKernelGGL(ozaki_int8_gemm<128>, dim3((unsigned)tiles), dim3(TNT), 0, st, Asl, Bsl,
                           rexp, cexp, C, m, n, L.Mp, L.Np, L.KTC, ldc, slices, nsteps,
                           accumulate ? 1 : 0);
    if ((rc = mgx::launch_status("ozaki_int8_gemm")) != MGX_OK) return rc;
    return mark(3);
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
