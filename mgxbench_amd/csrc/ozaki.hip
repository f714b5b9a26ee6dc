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
// One workgroup owns a TB x TB block of C.  TB = 256: 8 wave64s (2 x 4), two
// per SIMD, each holding a 128 x 64 INT32 accumulator tile (128 accumulation
// registers per lane); TB = 128: 4 waves (2 x 2) of 64 x 64, two workgroups per
// CU.  Two waves per SIMD let one wave's LDS reads, barrier wait and load
// issue overlap the other's MFMAs.
//
// Operand staging: a ring of NSTAGE LDS slots, each holding the A and B tiles
// of one 64-byte k stage, is filled by direct global->LDS loads
// (global_load_lds_dwordx4: lane l of a wave writes its 16 bytes at l*16 of a
// 1 KiB run, no registers involved).  Stage s+NSTAGE-1 is requested while
// stage s is multiplied, so NSTAGE-1 stages of loads are in flight.  Stages are
// published with an explicit vmcnt wait plus a plain s_barrier (a full
// __syncthreads would also drain the younger stages' loads).
// In LDS a tile row is 64 bytes; 16-byte chunk c of row r sits at position
// c ^ swz(r) (below), which keeps the fragment reads free of bank conflicts.
//
// Steps run in (group t, K chunk, pair p, k stage) order.  At the end of every
// K chunk the INT32 sums are folded into C in FP64:
//     x = ldexp((double)G, e_i + f_j + 4 - 8t)      (exact)
//     C = x  or  C_in + x   on the first fold (plain / accumulate form),
//     C = C + x             on every later fold.
// so C is read and written once per fold; C stays in HBM between folds.

constexpr int TGROUP_M = 8;       // C block rows per raster group
constexpr int NSTAGE = 4;         // LDS ring depth (stages)
constexpr int CHUNK_GRAIN = 2;    // K chunks are whole multiples of this many 64-k tiles

// Group-t accumulation chunk in 64-deep k tiles: the largest whole multiple
// of CHUNK_GRAIN tiles whose INT32 sums over all np pairs of the group stay
// below 2^31.
__device__ inline int chunk_tiles(int np) {
    int chunk = (int)max(1ll, 2147483647ll / (np * SLICE_MAX_PRODUCT * KT));
    return max(1, chunk / CHUNK_GRAIN) * CHUNK_GRAIN;
}

// Position of the next stage to load, walked in the product's step order
// (group t, K chunk, pair p, k tile).  The common advance adds one k tile to
// both operand pointers; group, chunk and pair changes take the slow path.
// Past the last step the cursor stays on the last stage, so the load that
// keeps the per-step load count uniform re-reads a valid tile into a slot no
// step reads again.
struct LoadCursor {
    const int8_t *ga, *gb;   // tile base of the current stage (no lane offset)
    int kt, c1;              // current k tile, end of the current chunk
    int t, p, pmax, c0, chunk;
    int s, ktiles;
    int64_t mkt, nkt;        // bytes per k tile of one A / B slice tile row band
    int64_t a_slice, b_slice;
    const int8_t *abase, *bbase;

    __device__ void place() {
        ga = abase + (int64_t)(p - 1) * a_slice + kt * mkt;
        gb = bbase + (int64_t)(t - p - 1) * b_slice + kt * nkt;
    }
    __device__ void begin_group() {
        p = max(1, t - s);
        pmax = min(s, t - 1);
        chunk = chunk_tiles(pmax - p + 1);
        c0 = 0;
        c1 = min(chunk, ktiles);
        kt = 0;
    }
    __device__ void start(int s_, int ktiles_, const int8_t *abase_, const int8_t *bbase_,
                          int64_t mkt_, int64_t nkt_, int64_t a_slice_, int64_t b_slice_) {
        s = s_;
        ktiles = ktiles_;
        abase = abase_;
        bbase = bbase_;
        mkt = mkt_;
        nkt = nkt_;
        a_slice = a_slice_;
        b_slice = b_slice_;
        t = 2;
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
            p = max(1, t - s);
            kt = c0;
        } else if (t <= s) {
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

// TB x TB block tile computed by WGM x WGN waves (wave tile TB/WGM x TB/WGN).
// The 256 tile runs 8 waves, two per SIMD, so one wave's LDS reads, barrier
// and load issue overlap the other wave's MFMAs.
template <int TB, int WGM, int WGN>
struct TileShape {
    static constexpr int NW = WGM * WGN;           // waves per workgroup
    static constexpr int THREADS = 64 * NW;
    static constexpr int WTM = TB / WGM;           // wave tile rows
    static constexpr int WTN = TB / WGN;           // wave tile columns
    static constexpr int FI = WTM / 16;            // 16x16 fragments along rows
    static constexpr int FJ = WTN / 16;            // ... along columns
    static constexpr int TILE = TB * KT;           // bytes of one operand tile per stage
    static constexpr int SLOT = 2 * TILE;          // A tile then B tile
    static constexpr int LPW = TILE / 1024 / NW;   // 1 KiB loads per wave per operand per stage
    static constexpr int FOLD_ROWS = 8;            // C rows per fold pass
    static constexpr int FOLD_STR = WTN + 4;       // ints per row of a wave's fold scratch
    static constexpr int FOLD_WAVE = FOLD_ROWS * FOLD_STR;
    static constexpr int LDS_BYTES = NSTAGE * SLOT + NW * FOLD_WAVE * 4;
    static_assert(WTN == 64, "the fold maps one C column per lane");
    static_assert(LPW >= 1 && LPW * NW * 1024 == TILE, "loads must cover the tile");
    static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");
};
using ShapeLarge = TileShape<256, 2, 4>;
using ShapeSmall = TileShape<128, 2, 2>;

template <int TB, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(64 * WGM * WGN, MINB) void ozaki_int8_gemm(
    const int8_t *__restrict__ Asl, const int8_t *__restrict__ Bsl,
    const int32_t *__restrict__ rexp, const int32_t *__restrict__ cexp, double *__restrict__ C,
    int64_t M, int64_t N, int64_t Mp, int64_t Np, int64_t KTC, int64_t ldc, int s, int accumulate) {
    using S = TileShape<TB, WGM, WGN>;
    constexpr int FI = S::FI, FJ = S::FJ, TILE = S::TILE, SLOT = S::SLOT, LPW = S::LPW;
    constexpr int FOLD_ROWS = S::FOLD_ROWS, FOLD_STR = S::FOLD_STR;
    // operand ring, then a private fold scratch region per wave
    __shared__ __attribute__((aligned(1024))) int8_t lds[S::LDS_BYTES];

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
        tm = first_m + (b % per_group) % gsize;
        tn = (b % per_group) / gsize;
    }
    const int64_t m0 = (int64_t)tm * TB;
    const int64_t n0 = (int64_t)tn * TB;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN;
    const int wn = wave % WGN;
    const int lk = lane >> 4;
    const int lc = lane & 15;

    // Direct-load assignment: load j of this wave fills tile rows
    // 16*(wave*LPW + j) .. +16; lane l reads row +l/4, 16-byte chunk
    // (l%4) ^ swz(l/4), landing at position l%4 of that row in LDS.
    // Consecutive loads are 1 KiB apart in HBM and in LDS.
    const int lane_off = (16 * wave * LPW + (lane >> 2)) * KT + (((lane & 3) ^ swz(lane >> 2)) << 4);

    v4i acc[FI][FJ];
    const v4i vzero = {0, 0, 0, 0};
    constexpr int LPS = 2 * LPW;   // load instructions per wave per stage

    const int ktiles = (int)KTC;
    LoadCursor cur;
    cur.start(s, ktiles, Asl + m0 * KT, Bsl + n0 * KT, Mp * KT, Np * KT, KTC * Mp * KT,
              KTC * Np * KT);
    const uint32_t voff = (uint32_t)lane_off;
    int8_t *const lds_w = lds + wave * LPW * 1024;   // this wave's load runs
    auto issue = [&](int slot) {
        int8_t *la = lds_w + slot * SLOT;
#pragma unroll
        for (int j = 0; j < LPW; ++j) {
            dma16(cur.ga + voff + j * 1024, la + j * 1024);
            dma16(cur.gb + voff + j * 1024, la + TILE + j * 1024);
        }
        cur.advance();
    };
    // prologue: stages 0 .. NSTAGE-2.  Every step then issues exactly one
    // stage, so NSTAGE-1 stages are in flight at each step's wait.
    for (int st = 0; st < NSTAGE - 1; ++st) issue(st);
    bool first_fold = true;
    int *fold_buf = (int *)(lds + NSTAGE * SLOT) + wave * S::FOLD_WAVE;
    int slot = 0;

    for (int t = 2; t <= s + 1; ++t) {
        const int np = min(s, t - 1) - max(1, t - s) + 1;
        const int chunk = chunk_tiles(np);
        for (int c0 = 0; c0 < ktiles; c0 += chunk) {
            const int nchunk = min(chunk, ktiles - c0) * np;   // steps in this chunk
#pragma unroll
            for (int i = 0; i < FI; ++i)
#pragma unroll
                for (int j = 0; j < FJ; ++j) acc[i][j] = vzero;
            for (int n = 0; n < nchunk; ++n) {
                // this wave's loads of the current stage are done once at most
                // the NSTAGE-2 younger stages are outstanding; its reads of the
                // previous stage completed before the previous MFMAs
                wait_vmcnt<(NSTAGE - 2) * LPS>();
                block_barrier();   // every wave's loads are in LDS; the previous slot is free
                issue((slot + NSTAGE - 1) & (NSTAGE - 1));

                const int8_t *la = lds + slot * SLOT;
                const int8_t *lb = la + TILE;
                v4i af[FI], bf[FJ];
#pragma unroll
                for (int i = 0; i < FI; ++i) {
                    const int r = wm * S::WTM + i * 16 + lc;
                    af[i] = *(const v4i *)(la + r * KT + ((lk ^ swz(r)) << 4));
                }
#pragma unroll
                for (int j = 0; j < FJ; ++j) {
                    const int r = wn * S::WTN + j * 16 + lc;
                    bf[j] = *(const v4i *)(lb + r * KT + ((lk ^ swz(r)) << 4));
                }
                // all fragment reads are issued before the first MFMA (the
                // scheduler would otherwise reuse one register per A fragment
                // and wait for each read in turn)
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < FI; ++i)
#pragma unroll
                    for (int j = 0; j < FJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
                slot = (slot + 1) & (NSTAGE - 1);
            }
            const int t_cur = t;

        // Fold: each wave dumps its INT32 accumulators to its private LDS
        // scratch, FOLD_ROWS C rows per pass (pass (i, h): fragment row i, rows
        // 8h..8h+7, held by the lanes with lk/2 == h), and folds them into C,
        // one column per lane.  A wave's LDS operations complete in order, so
        // no workgroup barrier is needed.
        {
            const int64_t fcol = n0 + wn * S::WTN + lane;
            const int ecol = fcol < N ? cexp[fcol] : 0;
#pragma unroll
            for (int i = 0; i < FI; ++i) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    // INT32 layout: lane l, register r holds row 4*(l/16)+r, column l%16.
                    if ((lk >> 1) == h) {
#pragma unroll
                        for (int j = 0; j < FJ; ++j)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                fold_buf[((lk & 1) * 4 + r) * FOLD_STR + j * 16 + lc] = acc[i][j][r];
                    }
                    asm volatile("" ::: "memory");
#pragma unroll 2
                    for (int rho = 0; rho < FOLD_ROWS; ++rho) {
                        const int64_t row = m0 + wm * S::WTM + i * 16 + h * 8 + rho;
                        if (row >= M) break;
                        const int er = rexp[row];
                        double *c_row = C + row * ldc;
                        if (fcol < N) {
                            const int g = fold_buf[rho * FOLD_STR + lane];
                            const double x = (er == EXP_BAD || ecol == EXP_BAD)
                                                 ? __longlong_as_double(0x7ff8000000000000ll)
                                                 : ldexp((double)g, er + ecol + 4 - 8 * t_cur);
                            c_row[fcol] = (first_fold && !accumulate) ? x : c_row[fcol] + x;
                        }
                    }
                    asm volatile("" ::: "memory");
                }
            }
            first_fold = false;
            // the fold's C traffic retires before the stage counting resumes
            wait_vmcnt<0>();
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
    if (tile == MGX_TILE_LARGE)
        hipLaunchKernelGGL((ozaki_int8_gemm<256, 2, 4, 1>), dim3((unsigned)tiles),
                           dim3(ShapeLarge::THREADS), 0, st, Asl, Bsl, rexp, cexp, C, m, n, L.Mp,
                           L.Np, L.KTC, ldc, slices, accumulate ? 1 : 0);
    else
        hipLaunchKernelGGL((ozaki_int8_gemm<128, 2, 2, 2>), dim3((unsigned)tiles),
                           dim3(ShapeSmall::THREADS), 0, st, Asl, Bsl, rexp, cexp, C, m, n, L.Mp,
                           L.Np, L.KTC, ldc, slices, accumulate ? 1 : 0);
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
