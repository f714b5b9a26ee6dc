// Plain FP64 GEMM on the CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4_f64),
// and its panel-fused form.
//
// Replaces the reference's tiled kernels sharedABMultiply /
// sharedABMultiplyAccumulate (1D-row-block/mulmat_kernel.cu:39-80, 100-142;
// 2D-block-matrix/mulmat_kernel.cu:36-75, 94-136), their launchers
// ABMultiply / ABMultiplyAccumulateAsync (1D-row-block/mulmat_kernel.cu:86-92,
// 148-158), and the per-panel launch sequence of the SUMMA loop
// (2D-block-matrix/mmb.cu:239-361).  Semantics kept: row-major operands with
// leading dimensions, out-of-range A columns contribute nothing, and every
// panel's partial product is formed from zero and then added into C once, in
// panel order (C += A_p . B_p per panel, as the reference's one accumulate
// launch per panel does).
//
// One kernel, dgemm_panels, does all of it: C (+)= sum_p A[:, p] . B[p, :]
// over panels of width `panel` (ragged last panel).  A plain product is one
// panel of depth K.  Each workgroup holds its C tile in registers for the
// whole reduction: the running sum and the current panel's MFMA partial are
// two register sets, and at a panel's end the partial is added into the sum
// (one FP64 add per element) and reset, so C is read and written once per
// launch however many panels the launch covers.
//
// Design (gfx950):
//  * Tiles: panel-fused, 128 x 128 C block per workgroup of 8 wave64s
//    (2 x 4), each wave 64 x 32 as 4 x 2 MFMA fragments (two register sets:
//    running sum and partial); one workgroup per CU, two waves per SIMD.
//    Small variant: 64 x 64 with 4 waves of 32 x 32, two per CU.  A
//    one-panel product (the partial is the result) runs a 256 x 128 block,
//    8 waves of 64 x 64, with a 3-deep ring (fewer loads per MFMA).
//  * Staging: a ring of NSTAGE LDS slots, each the A and B tiles of one
//    16-deep k stage (32 KiB at 128 x 128), filled by direct global->LDS loads
//    (global_load_lds_dwordx4: lane l writes its 16 bytes at l * 16 of a 1 KiB
//    run, no registers involved); every stage issues one stage NSTAGE-1
//    ahead, published with an explicit vmcnt wait plus a raw s_barrier.
//  * A stays row-major in LDS (128-byte rows of 16 k).  16-byte chunk c of
//    row r sits at position c ^ (r % 8): the loads pick which chunk each lane
//    fetches, so the swizzle costs nothing.  An A fragment read
//    (ds_read_b128) gives a lane two consecutive k of its row, which feed two
//    consecutive MFMA steps: within a pair of steps lane group g = l / 16
//    takes k = 8P + 2g + h for step h of pair P.  B is read with the same k
//    assignment; the 8 lanes of each read pass touch 8 different chunk
//    positions, so neither operand has bank conflicts.
//  * Issue order: each pair of k steps issues MFMA_LEAD MFMAs, then the
//    fragment reads of the next pair, then its other MFMAs, so the reads
//    overlap matrix work without a wait in front of the first MFMA; the
//    B fragments of a pair are separate registers per step, which keeps the
//    compiler from inserting copies.
//  * Edges: rows of A past M and columns of B past N are clamped to valid
//    addresses and their outputs discarded.  A stage that crosses the end of
//    its panel, and every stage of operands that are not 16-byte aligned with
//    even leading dimensions and N, is loaded through registers with explicit
//    bounds (zeros past the panel end) and stored to the same LDS layout.
//  * Tile order: persistent grid of one workgroup per CU; the tiles of the
//    grouped raster (8 tile rows per group) are dealt out in 8 contiguous
//    ranges, one per XCD (workgroup w runs on XCD w % 8), so the tiles in
//    flight on an XCD share A rows and B columns in its L2.
#include "mgx_internal.hpp"

#include <mutex>
#include <type_traits>

namespace {

constexpr int GROUP_M = 8;    // C tile rows per raster group

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// s_waitcnt with vmcnt <= N, other counters untouched (gfx9 encoding).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// s_waitcnt lgkmcnt(0), other counters untouched.
__device__ inline void wait_lgkm0() {
    __builtin_amdgcn_s_waitcnt(0xF | (0x7 << 4) | (0x0 << 8) | (0x3 << 14));
}

__device__ inline void block_barrier() {
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// 16-byte direct global->LDS load: lane l writes its bytes at lds_run + 16*l
// (lds_run must be wave-uniform).
__device__ inline void dma16(const double *g, char *lds_run) {
    __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void *)lds_run, 16, 0,
                                     0);
}

// A pointer the caller knows to be wave-uniform, moved to scalar registers
// (lets a load use the scalar-base + 32-bit vector-offset address form).
__device__ inline const char *uniform_ptr(const char *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const char *)(((uint64_t)hi << 32) | lo);
}

// BM x BN C tile per workgroup of WGM x WGN waves; k stages of KS in a ring
// of NSTAGE LDS slots.
template <int BM_, int BN_, int WGM_, int WGN_, int KS_, int NSTAGE_>
struct DShape {
    static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, KS = KS_, NSTAGE = NSTAGE_;
    static constexpr int NW = WGM * WGN;
    static constexpr int THREADS = 64 * NW;
    static constexpr int WTM = BM / WGM;          // wave tile rows
    static constexpr int WTN = BN / WGN;          // wave tile columns
    static constexpr int FI = WTM / 16;           // MFMA fragments along rows
    static constexpr int FJ = WTN / 16;           // ... along columns
    static constexpr int NP = KS / 8;             // pairs of MFMA steps per stage
    static constexpr int CPR = KS / 2;            // 16-byte chunks per A row of a stage
    static constexpr int RPR = 64 / CPR;          // A rows per 1 KiB run
    static constexpr int A_BYTES = BM * KS * 8;   // A tile of one stage
    static constexpr int B_BYTES = KS * BN * 8;   // B tile of one stage
    static constexpr int SLOT = A_BYTES + B_BYTES;
    static constexpr int RA = A_BYTES / 1024 / NW;   // 1 KiB A runs per wave per stage
    static constexpr int RB = B_BYTES / 1024 / NW;   // ... B runs
    static constexpr int LPB = BN / 2;            // lanes per B k-row (16 bytes each)
    static constexpr int BROWS = 64 / LPB;        // B k-rows per run
    static constexpr int LDS_BYTES = NSTAGE * SLOT;
    static constexpr int LPS = RA + RB;           // direct loads per wave per stage
    static_assert(KS % 16 == 0 && CPR >= 8 && RPR * CPR == 64, "stage depth");
    static_assert(RA >= 1 && RA * NW * 1024 == A_BYTES, "A runs");
    static_assert(RB >= 1 && RB * NW * 1024 == B_BYTES, "B runs");
    static_assert(BROWS >= 1 && BROWS * LPB == 64, "B run shape");
    static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");
    static_assert((NSTAGE - 2) * LPS < 64 && NSTAGE >= 2, "ring depth");
};
// MFMAs of a pair of k steps issued before the next pair's fragment reads
// (so no LDS read stands in front of the matrix core), and MFMAs of a
// stage's last pair issued between the stage's publish wait and its next
// load issue.
constexpr int MFMA_LEAD = 2;
constexpr int MFMA_BEFORE_ISSUE = 2;
// panel-fused tiles (running sum + partial in registers)
using Large = DShape<128, 128, 2, 4, 16, 4>;
using Small = DShape<64, 64, 2, 2, 16, 4>;
// one-panel tile (the partial is the result): 256 x 128, 8 waves of 64 x 64
using Wide = DShape<256, 128, 4, 2, 16, 3>;

struct PanelArgs {
    const double *A, *B;
    double *C;
    int64_t M, N, K, lda, ldb, ldc;
    int64_t panel;        // panel width; K for a plain product
    int64_t stages;       // k stages over all panels
    int accumulate;       // C += (else C =)
    int dma;              // operands allow direct 16-byte loads
    int tiles_m, tiles_n;
};

// Walks the k stages panel by panel: stage [k0, k0 + KS) of the panel that
// ends at pend.  Past the last stage it stays there.
template <int KS>
struct StageCursor {
    int K, panel, pend, k0;
    __device__ void start(int K_, int panel_) {
        K = K_;
        panel = panel_;
        k0 = 0;
        pend = panel < K ? panel : K;
    }
    __device__ bool full() const { return k0 + KS <= pend; }
    __device__ bool last_in_panel() const { return k0 + KS >= pend; }
    __device__ void advance() {
        if (k0 + KS < pend) {
            k0 += KS;
        } else if (pend < K) {
            k0 = pend;
            pend = pend + panel < K ? pend + panel : K;
        }
    }
};

// Raster index -> (tile_m, tile_n): groups of GROUP_M tile rows, column-major
// within a group.
__device__ inline void raster(int b, int tiles_m, int tiles_n, int &tm, int &tn) {
    const int per_group = GROUP_M * tiles_n;
    const int g = b / per_group;
    const int first_m = g * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int r = b - g * per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// FUSED: the running sum of the C tile and the current panel's partial are
// two register sets (any number of panels).  Otherwise the launch is one
// panel and the partial itself is added to (or written as) C at the end,
// which frees the registers for a larger tile; the result is the same.
template <class S, bool FUSED>
__global__ __launch_bounds__(S::THREADS, 1) void dgemm_panels(PanelArgs a) {
    constexpr int FI = S::FI, FJ = S::FJ, SLOT = S::SLOT, NSTAGE = S::NSTAGE, KS = S::KS;
    __shared__ __attribute__((aligned(1024))) char lds[S::LDS_BYTES];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / S::WGN;
    const int wn = wave % S::WGN;
    const int64_t M = a.M, N = a.N;
    const int tiles = a.tiles_m * a.tiles_n;
    const int G = gridDim.x;
    // Tiles of workgroup w: with one round of workgroups, tile w (renumbered
    // so that each XCD takes a contiguous range); with fewer workgroups than
    // tiles, every (G/8)-th tile of XCD w % 8's contiguous range.
    const int w = blockIdx.x;
    int first, stride, limit;
    if (G >= tiles) {
        first = ((tiles & 7) == 0) ? (w & 7) * (tiles >> 3) + (w >> 3) : w;
        stride = tiles;
        limit = tiles;
    } else {
        const int per_xcd = (tiles + 7) >> 3;
        first = (w & 7) * per_xcd + (w >> 3);
        stride = G >> 3;
        limit = min(((w & 7) + 1) * per_xcd, tiles);
    }

    // lane-constant pieces of the load and read addresses
    const int a_row_in_run = lane / S::CPR;                // A: row within a run
    const int a_pos = lane % S::CPR;                       // A: LDS chunk position of the lane
    const int b_row_in_run = lane / S::LPB;                // B: k-row within a run
    const int b_col = 2 * (lane % S::LPB);                 // B: first column of the lane
    const int lc = lane & 15, lg = lane >> 4;
    const v4d vzero = {0.0, 0.0, 0.0, 0.0};

    for (int t = first; t < limit; t += stride) {
        int tm, tn;
        raster(t, a.tiles_m, a.tiles_n, tm, tn);
        const int64_t m0 = (int64_t)tm * S::BM;
        const int64_t n0 = (int64_t)tn * S::BN;

        // running sum of the C tile (FUSED), and the current panel's partial
        v4d csum[FUSED ? FI : 1][FUSED ? FJ : 1], acc[FI][FJ];
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
            for (int j = 0; j < FJ; ++j) {
                acc[i][j] = vzero;
                if constexpr (FUSED) {
                    csum[i][j] = vzero;
                    if (a.accumulate) {
                        // FP64 MFMA layout: lane l, register r -> row 4r + l/16, column l%16
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int64_t row = m0 + wm * S::WTM + 16 * i + 4 * r + lg;
                            const int64_t col = n0 + wn * S::WTN + 16 * j + lc;
                            csum[i][j][r] = (row < M && col < N) ? a.C[row * a.ldc + col] : 0.0;
                        }
                    }
                }
            }

        // per-lane global rows / columns of the direct loads (clamped inside
        // the matrix; outputs of clamped rows and columns are discarded).
        // Chunk c of A row r lands at LDS position c ^ (r % 8): the lane at
        // position p fetches chunk p ^ (r % 8).
        const double *a_src[S::RA];
        int a_chunk[S::RA];
        uint32_t a_off[S::RA];    // byte offsets from the tile's A rows (direct loads)
#pragma unroll
        for (int j = 0; j < S::RA; ++j) {
            const int rr = (wave * S::RA + j) * S::RPR + a_row_in_run;
            a_chunk[j] = a_pos ^ (rr & 7);
            const int64_t row = m0 + rr;
            a_src[j] = a.A + row * a.lda + 2 * a_chunk[j];
            const int64_t crow = (row < M ? row : M - 1) - m0;
            a_off[j] = (uint32_t)((crow * a.lda + 2 * a_chunk[j]) * 8);
        }
        int64_t b_gcol = n0 + b_col;
        if (b_gcol > N - 2) b_gcol = N - 2 >= 0 ? N - 2 : 0;
        uint32_t b_off[S::RB];    // byte offsets from the stage's B rows at column n0
#pragma unroll
        for (int j = 0; j < S::RB; ++j) {
            const int krow = (wave * S::RB + j) * S::BROWS + b_row_in_run;
            b_off[j] = (uint32_t)((krow * a.ldb + (b_gcol - n0)) * 8);
        }
        const char *a_tile = (const char *)(a.A + m0 * a.lda);
        const char *b_tile = (const char *)(a.B + n0);

        StageCursor<KS> ld;
        ld.start((int)a.K, (int)a.panel);
        // Issues the loads of the loader's stage into `slot`.
        auto issue = [&](int slot) {
            char *base = lds + slot * SLOT;
            if (a.dma && ld.full()) {
                // wave-uniform stage bases plus per-lane 32-bit offsets
                const char *ab = uniform_ptr(a_tile + (int64_t)ld.k0 * 8);
                const char *bb = uniform_ptr(b_tile + (int64_t)ld.k0 * a.ldb * 8);
#pragma unroll
                for (int j = 0; j < S::RA; ++j)
                    dma16((const double *)(ab + a_off[j]), base + (wave * S::RA + j) * 1024);
#pragma unroll
                for (int j = 0; j < S::RB; ++j)
                    dma16((const double *)(bb + b_off[j]),
                          base + S::A_BYTES + (wave * S::RB + j) * 1024);
            } else {
                // through registers, zero past the panel end or the matrix
                v2d va[S::RA], vb[S::RB];
#pragma unroll
                for (int j = 0; j < S::RA; ++j) {
                    const int64_t row = m0 + (wave * S::RA + j) * S::RPR + a_row_in_run;
                    const int k = ld.k0 + 2 * a_chunk[j];
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        va[j][e] = (row < M && k + e < ld.pend) ? a_src[j][ld.k0 + e] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < S::RB; ++j) {
                    const int64_t krow = (int64_t)ld.k0 + (wave * S::RB + j) * S::BROWS +
                                         b_row_in_run;
                    const int64_t col = n0 + b_col;
#pragma unroll
                    for (int e = 0; e < 2; ++e)
                        vb[j][e] = (krow < ld.pend && col + e < N) ? a.B[krow * a.ldb + col + e]
                                                                   : 0.0;
                }
#pragma unroll
                for (int j = 0; j < S::RA; ++j)
                    *(v2d *)(base + (wave * S::RA + j) * 1024 + lane * 16) = va[j];
#pragma unroll
                for (int j = 0; j < S::RB; ++j)
                    *(v2d *)(base + S::A_BYTES + (wave * S::RB + j) * 1024 + lane * 16) = vb[j];
                wait_lgkm0();
            }
            ld.advance();
        };

        const int64_t T = a.stages;
        if (T > 0) {
            for (int st = 0; st < NSTAGE - 1; ++st) issue(st);
            int pslot = NSTAGE - 1;
            // the next stage's loads are in LDS once at most the NSTAGE-2
            // younger stages are outstanding; after the barrier every wave's
            // are, and every wave is done reading the stage before it, whose
            // slot then takes the stage NSTAGE-1 ahead
            auto publish_wait = [&]() {
                wait_vmcnt<(NSTAGE - 2) * S::LPS>();
                block_barrier();
            };
            auto publish_issue = [&]() {
                const int freed = pslot;
                pslot = pslot + 1 == NSTAGE ? 0 : pslot + 1;
                issue(freed);
            };
            // fragments of pair P (k 8P .. 8P+7 of the stage) from `slot`: A as
            // one 16-byte read per fragment (k and k+1 of the lane's row), B
            // as single doubles (k row and k+1 row), kept apart so the reads
            // can be merged freely
            auto read = [&](int slot, int P, v2d (&fa)[FI], double (&fb)[2][FJ]) {
                const char *la = lds + slot * SLOT;
                const char *lb = la + S::A_BYTES;
#pragma unroll
                for (int i = 0; i < FI; ++i) {
                    const int r = wm * S::WTM + 16 * i + lc;
                    const int pos = (4 * P + lg) ^ (r & 7);
                    fa[i] = *(const v2d *)(la + r * (KS * 8) + pos * 16);
                }
#pragma unroll
                for (int j = 0; j < FJ; ++j) {
                    const int col = wn * S::WTN + 16 * j + lc;
                    const int k = 8 * P + 2 * lg;
                    fb[0][j] = *(const double *)(lb + k * (S::BN * 8) + col * 8);
                    fb[1][j] = *(const double *)(lb + (k + 1) * (S::BN * 8) + col * 8);
                }
            };
            // MFMAs [q0, q1) of the pair's 2 * FI * FJ, in (h, i, j) order
            auto mfma = [&](const v2d (&fa)[FI], const double (&fb)[2][FJ], auto q0, auto q1) {
                constexpr int Q0 = decltype(q0)::value, Q1 = decltype(q1)::value;
#pragma unroll
                for (int q = Q0; q < Q1; ++q) {
                    const int h = q / (FI * FJ), i = (q / FJ) % FI, j = q % FJ;
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i][h], fb[h][j],
                                                                     acc[i][j], 0, 0, 0);
                }
            };
            using Q0 = std::integral_constant<int, 0>;
            using QH = std::integral_constant<int, MFMA_BEFORE_ISSUE>;
            using QA = std::integral_constant<int, MFMA_LEAD>;
            using QE = std::integral_constant<int, 2 * FI * FJ>;
            StageCursor<KS> cs;
            cs.start((int)a.K, (int)a.panel);
            v2d fa[2][FI];
            double fb[2][2][FJ];
            publish_wait();
            publish_issue();
            read(pslot, 0, fa[0], fb[0]);
            wait_lgkm0();
            for (int64_t s = 0; s < T; ++s) {
#pragma unroll
                for (int P = 0; P < S::NP; ++P) {
                    const int cur = P & 1;
                    if (P + 1 < S::NP) {
                        // a few MFMAs first, so no fragment read issued just
                        // before them stands in front of the matrix core
                        mfma(fa[cur], fb[cur], Q0{}, QA{});
                        __builtin_amdgcn_sched_barrier(0);
                        read(pslot, P + 1, fa[cur ^ 1], fb[cur ^ 1]);
                        __builtin_amdgcn_sched_barrier(0);
                        mfma(fa[cur], fb[cur], QA{}, QE{});
                        __builtin_amdgcn_sched_barrier(0);
                    } else {
                        wait_lgkm0();   // the stage's fragments are in registers
                        const bool more = s + 1 < T;
                        if (more) publish_wait();
                        __builtin_amdgcn_sched_barrier(0);
                        mfma(fa[cur], fb[cur], Q0{}, QH{});
                        __builtin_amdgcn_sched_barrier(0);
                        if (more) publish_issue();
                        // unconditional, so the fragment registers keep one
                        // role around the loop; after the last stage it
                        // re-reads the current slot and the values go unused
                        read(pslot, 0, fa[0], fb[0]);
                        __builtin_amdgcn_sched_barrier(0);
                        mfma(fa[cur], fb[cur], QH{}, QE{});
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                if constexpr (FUSED) {
                    if (cs.last_in_panel()) {
                        // the panel's partial joins the running sum: C += A_p . B_p
#pragma unroll
                        for (int i = 0; i < FI; ++i)
#pragma unroll
                            for (int j = 0; j < FJ; ++j) {
                                csum[i][j] += acc[i][j];
                                acc[i][j] = vzero;
                            }
                    }
                    cs.advance();
                }
            }
        }

#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
            for (int j = 0; j < FJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t row = m0 + wm * S::WTM + 16 * i + 4 * r + lg;
                    const int64_t col = n0 + wn * S::WTN + 16 * j + lc;
                    if (row < M && col < N) {
                        double *c = a.C + row * a.ldc + col;
                        if constexpr (FUSED)
                            *c = csum[i][j][r];
                        else
                            *c = a.accumulate ? *c + acc[i][j][r] : 0.0 + acc[i][j][r];
                    }
                }
        // loads past the last stage and the C traffic retire, and every wave
        // is done with the ring, before the next tile reuses it
        wait_vmcnt<0>();
        block_barrier();
    }
}

int cu_count() {
    static std::mutex mu;
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

template <class S, bool FUSED>
int launch_panels(const double *A, const double *B, double *C, int64_t m, int64_t n, int64_t k,
                  int64_t panel, int64_t lda, int64_t ldb, int64_t ldc, bool accumulate,
                  hipStream_t st) {
    const int64_t tm = (m + S::BM - 1) / S::BM, tn = (n + S::BN - 1) / S::BN;
    MGX_ARG_CHECK(k <= (1ll << 30), "k too large for one launch (at most 2^30)");
    MGX_ARG_CHECK(tm * tn <= 0x7fffffffll && tm <= 0x7fffffff && tn <= 0x7fffffff,
                  "matrix too large for one launch");
    PanelArgs a;
    a.A = A;
    a.B = B;
    a.C = C;
    a.M = m;
    a.N = n;
    a.K = k;
    a.lda = lda;
    a.ldb = ldb;
    a.ldc = ldc;
    a.panel = panel < k ? panel : (k > 0 ? k : 1);
    // k stages: whole panels of ceil(panel / KS) stages, then the ragged one
    const int64_t spp = (a.panel + S::KS - 1) / S::KS;
    a.stages = k > 0 ? (k / a.panel) * spp + ((k % a.panel) + S::KS - 1) / S::KS : 0;
    a.accumulate = accumulate ? 1 : 0;
    // direct loads need 16-byte aligned A and B rows and stage starts: even
    // leading dimensions, even N (the last B pair of a row stays in the row)
    // and even panel starts
    a.dma = (((uintptr_t)A | (uintptr_t)B) & 15) == 0 && (lda & 1) == 0 && (ldb & 1) == 0 &&
            (n & 1) == 0 && ((a.panel & 1) == 0 || a.panel >= k) &&
            // per-lane byte offsets within a tile fit 32 bits
            (int64_t)S::BM * lda * 8 < (1ll << 31) && (int64_t)S::KS * ldb * 8 < (1ll << 31);
    a.tiles_m = (int)tm;
    a.tiles_n = (int)tn;
    const int64_t tiles = tm * tn;
    // one round of workgroups when it fits the CUs, else a persistent grid
    const int per_cu = (S::LDS_BYTES > 80 * 1024) ? 1 : 2;
    const int64_t slots = (int64_t)cu_count() * per_cu;
    const int64_t grid = tiles <= slots ? tiles : (slots / 8) * 8;
    hipLaunchKernelGGL((dgemm_panels<S, FUSED>), dim3((unsigned)grid), dim3(S::THREADS), 0, st, a);
    return mgx::launch_status("dgemm_panels");
}

int launch_dgemm(int tile, bool accumulate, const double *A, const double *B, double *C,
                 int64_t m, int64_t n, int64_t k, int64_t panel, int64_t lda, int64_t ldb,
                 int64_t ldc, hipStream_t st) {
    if (tile == MGX_TILE_DEFAULT) tile = MGX_TILE_LARGE;
    const bool one_panel = panel >= k;
    switch (tile) {
        case MGX_TILE_LARGE:
            // a one-panel product needs no running sum: the wider tile
            if (one_panel)
                return launch_panels<Wide, false>(A, B, C, m, n, k, panel, lda, ldb, ldc,
                                                  accumulate, st);
            return launch_panels<Large, true>(A, B, C, m, n, k, panel, lda, ldb, ldc, accumulate,
                                              st);
        case MGX_TILE_SMALL:
            return launch_panels<Small, true>(A, B, C, m, n, k, panel, lda, ldb, ldc, accumulate,
                                              st);
        default:
            return mgx::set_error(MGX_ERR_UNSUPPORTED, "unknown tile");
    }
}

}  // namespace

static int dgemm_entry(int tile, const double *A, const double *B, double *C, int64_t m,
                       int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc, int w,
                       void *stream) {
    mgx::clear_error();
    MGX_ARG_CHECK(w > 0, "tile width w must be positive");
    MGX_ARG_CHECK(m >= 0 && n >= 0 && k >= 0, "negative matrix dimension");
    if (m == 0 || n == 0) return MGX_OK;
    MGX_ARG_CHECK(C != nullptr, "C is NULL");
    MGX_ARG_CHECK(ldc >= n, "ldc < n");
    if (k > 0) {
        MGX_ARG_CHECK(A != nullptr && B != nullptr, "A or B is NULL");
        MGX_ARG_CHECK(lda >= k && ldb >= n, "lda < k or ldb < n");
    }
    return launch_dgemm(tile, false, A, B, C, m, n, k, k, lda, ldb, ldc, (hipStream_t)stream);
}

static int dgemm_accumulate_entry(int tile, const double *A, const double *B_chunk, double *C,
                                  int64_t m, int64_t n, int64_t k_total, int64_t k_offset,
                                  int64_t k_chunk, int64_t lda, int64_t ldb, int64_t ldc, int w,
                                  void *stream) {
    mgx::clear_error();
    MGX_ARG_CHECK(w > 0, "tile width w must be positive");
    MGX_ARG_CHECK(m >= 0 && n >= 0 && k_total >= 0 && k_offset >= 0 && k_chunk >= 0,
                  "negative matrix dimension or offset");
    if (m == 0 || n == 0) return MGX_OK;
    MGX_ARG_CHECK(C != nullptr, "C is NULL");
    MGX_ARG_CHECK(ldc >= n, "ldc < n");
    // Effective depth: A columns at or beyond k_total contribute nothing.
    int64_t k_eff = k_total - k_offset;
    if (k_eff > k_chunk) k_eff = k_chunk;
    if (k_eff < 0) k_eff = 0;
    if (k_eff > 0) {
        MGX_ARG_CHECK(A != nullptr && B_chunk != nullptr, "A or B_chunk is NULL");
        MGX_ARG_CHECK(lda >= k_offset + k_eff, "lda too small for k_offset + k_chunk");
        MGX_ARG_CHECK(ldb >= n, "ldb < n");
    }
    return launch_dgemm(tile, true, k_eff > 0 ? A + k_offset : A, B_chunk, C, m, n, k_eff, k_eff,
                        lda, ldb, ldc, (hipStream_t)stream);
}

extern "C" int mgx_dgemm(const double *A, const double *B, double *C, int64_t m, int64_t n,
                         int64_t k, int64_t lda, int64_t ldb, int64_t ldc, int w, void *stream) {
    return dgemm_entry(MGX_TILE_DEFAULT, A, B, C, m, n, k, lda, ldb, ldc, w, stream);
}

extern "C" int mgx_dgemm_accumulate(const double *A, const double *B_chunk, double *C,
                                    int64_t m, int64_t n, int64_t k_total, int64_t k_offset,
                                    int64_t k_chunk, int64_t lda, int64_t ldb, int64_t ldc,
                                    int w, void *stream) {
    return dgemm_accumulate_entry(MGX_TILE_DEFAULT, A, B_chunk, C, m, n, k_total, k_offset,
                                  k_chunk, lda, ldb, ldc, w, stream);
}

extern "C" int mgx_dgemm_ex(int tile, int accumulate, const double *A, const double *B, double *C,
                            int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                            void *stream) {
    if (accumulate)
        return dgemm_accumulate_entry(tile, A, B, C, m, n, k, 0, k, lda, ldb, ldc, 32, stream);
    return dgemm_entry(tile, A, B, C, m, n, k, lda, ldb, ldc, 32, stream);
}

extern "C" int mgx_dgemm_panels(int tile, int accumulate, const double *A, const double *B,
                                double *C, int64_t m, int64_t n, int64_t k, int64_t panel,
                                int64_t lda, int64_t ldb, int64_t ldc, void *stream) {
    mgx::clear_error();
    MGX_ARG_CHECK(m >= 0 && n >= 0 && k >= 0, "negative matrix dimension");
    MGX_ARG_CHECK(panel > 0, "panel width must be positive");
    if (m == 0 || n == 0) return MGX_OK;
    MGX_ARG_CHECK(C != nullptr, "C is NULL");
    MGX_ARG_CHECK(ldc >= n, "ldc < n");
    if (k > 0) {
        MGX_ARG_CHECK(A != nullptr && B != nullptr, "A or B is NULL");
        MGX_ARG_CHECK(lda >= k && ldb >= n, "lda < k or ldb < n");
    }
    return launch_dgemm(tile, accumulate != 0, A, B, C, m, n, k, panel, lda, ldb, ldc,
                        (hipStream_t)stream);
}
