// Plain FP64 GEMM on the CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the reference's tiled kernels sharedABMultiply /
// sharedABMultiplyAccumulate (1D-row-block/mulmat_kernel.cu:39-80, 100-142;
// 2D-block-matrix/mulmat_kernel.cu:36-75, 94-136) and their launchers
// ABMultiply / ABMultiplyAccumulateAsync (1D-row-block/mulmat_kernel.cu:86-92,
// 148-158).  Semantics kept: row-major operands with leading dimensions,
// zero contribution from out-of-range A columns (k_offset + kk >= k_total),
// and the accumulate form adds one per-chunk partial product into C.
//
// Design (gfx950): one workgroup of 256 lanes (4 wave64s) owns a 128x128
// block of C (64x64 for the small-tile variant); each wave owns a quarter,
// held as 16x16 MFMA accumulators (64 doubles per lane at 128x128).  A and B k-slabs of depth 16 are
// staged through LDS with two buffers: the global loads of slab t+1 are in
// flight while the MFMAs of slab t run, so one barrier per slab suffices.
// A is stored k-major in LDS so that both MFMA operands are read as 16
// consecutive doubles per 16-lane group.  Workgroup ids are remapped so
// that the tiles resident on one XCD (ids are dealt round-robin over the 8
// XCDs) form a compact group of C rows, which keeps their A and B slabs in
// that XCD's L2.
#include "mgx_internal.hpp"

namespace {

constexpr int BK = 16;
constexpr int NT = 256;
constexpr int GROUP_M = 8;   // C tile rows per raster group

typedef double v4d __attribute__((ext_vector_type(4)));

// Maps the linear workgroup id to a (tile_m, tile_n) pair: XCD-contiguous
// renumbering followed by a grouped raster over GROUP_M tile rows.
__device__ inline void tile_coords(int tiles_m, int tiles_n, int &tm, int &tn) {
    const int nb = tiles_m * tiles_n;
    int b = blockIdx.x;
    if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
    const int per_group = GROUP_M * tiles_n;
    const int g = b / per_group;
    const int first_m = g * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int r = b - g * per_group;
    tm = first_m + r % gsize;
    tn = r / gsize;
}

// BM x BN block of C per workgroup, 2x2 waves of (BM/2) x (BN/2), each held as
// (BM/32) x (BN/32) MFMA 16x16 accumulators.
template <int BM, int BN, bool ACCUM>
__global__ __launch_bounds__(NT, 2) void dgemm_mfma(
    const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C,
    int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc) {
    constexpr int SA = BM + 16;  // LDS row stride (doubles) of the k-major A slab
    constexpr int SB = BN + 16;  // LDS row stride (doubles) of the B slab
    constexpr int FI = BM / 32;  // MFMA tiles per wave along m
    constexpr int FJ = BN / 32;  // MFMA tiles per wave along n
    constexpr int EA = BM * BK / NT;  // A elements per lane per slab
    constexpr int EB = BK * BN / NT;  // B elements per lane per slab
    __shared__ double As[2][BK][SA];
    __shared__ double Bs[2][BK][SB];

    const int tiles_m = (int)((M + BM - 1) / BM);
    const int tiles_n = (int)((N + BN - 1) / BN);
    int tm, tn;
    tile_coords(tiles_m, tiles_n, tm, tn);
    const int64_t m0 = (int64_t)tm * BM;
    const int64_t n0 = (int64_t)tn * BN;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;

    // Global->LDS load assignment: each lane takes EA consecutive k of one A
    // row and EB consecutive columns of one B row.
    const int a_row = tid / (BK / EA);
    const int a_k = (tid % (BK / EA)) * EA;
    const int b_row = tid / (BN / EB);
    const int b_col = (tid % (BN / EB)) * EB;

    const int64_t ga_row = m0 + a_row;
    const bool a_row_ok = ga_row < M;
    const double *a_ptr = A + (a_row_ok ? ga_row : 0) * lda;
    const int64_t gb_col = n0 + b_col;

    double ra[EA], rb[EB];

    auto load_slab = [&](int64_t k0) {
#pragma unroll
        for (int j = 0; j < EA; ++j) {
            const int64_t kk = k0 + a_k + j;
            ra[j] = (a_row_ok && kk < K) ? a_ptr[kk] : 0.0;
        }
        const int64_t kb = k0 + b_row;
        const bool b_row_ok = kb < K;
        const double *b_ptr = B + (b_row_ok ? kb : 0) * ldb;
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            const int64_t c = gb_col + j;
            rb[j] = (b_row_ok && c < N) ? b_ptr[c] : 0.0;
        }
    };
    auto store_slab = [&](int buf) {
#pragma unroll
        for (int j = 0; j < EA; ++j) As[buf][a_k + j][a_row] = ra[j];
#pragma unroll
        for (int j = 0; j < EB; ++j) Bs[buf][b_row][b_col + j] = rb[j];
    };

    v4d acc[FI][FJ];
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j) acc[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};

    const int nslabs = (int)((K + BK - 1) / BK);
    const int lk = lane >> 4;   // k row within an MFMA step
    const int lc = lane & 15;   // row (A) / column (B) within a 16-wide fragment

    if (nslabs > 0) {
        load_slab(0);
        store_slab(0);
        __syncthreads();
    }
    for (int s = 0; s < nslabs; ++s) {
        const int buf = s & 1;
        const bool more = (s + 1) < nslabs;
        if (more) load_slab((int64_t)(s + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[FI], bf[FJ];
#pragma unroll
            for (int i = 0; i < FI; ++i) af[i] = As[buf][kk + lk][wm * (BM / 2) + i * 16 + lc];
#pragma unroll
            for (int j = 0; j < FJ; ++j) bf[j] = Bs[buf][kk + lk][wn * (BN / 2) + j * 16 + lc];
#pragma unroll
            for (int i = 0; i < FI; ++i)
#pragma unroll
                for (int j = 0; j < FJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (more) store_slab(buf ^ 1);
        __syncthreads();
    }

    // FP64 16x16 accumulator layout: lane l, register r holds row 4*r + l/16,
    // column l%16 of the fragment (register r covers rows 4r..4r+3).
#pragma unroll
    for (int i = 0; i < FI; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = m0 + wm * (BM / 2) + i * 16 + 4 * r + lk;
            if (row >= M) continue;
            double *c_row = C + row * ldc;
#pragma unroll
            for (int j = 0; j < FJ; ++j) {
                const int64_t col = n0 + wn * (BN / 2) + j * 16 + lc;
                if (col < N) {
                    if (ACCUM)
                        c_row[col] = c_row[col] + acc[i][j][r];
                    else
                        c_row[col] = acc[i][j][r];
                }
            }
        }
    }
}

template <int BM, int BN>
int launch_tile(bool accumulate, const double *A, const double *B, double *C, int64_t m,
                int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc, hipStream_t st) {
    const int64_t tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    MGX_ARG_CHECK(tiles <= 0x7fffffffll, "matrix too large for one launch");
    dim3 grid((unsigned)tiles), block(NT);
    if (accumulate)
        hipLaunchKernelGGL((dgemm_mfma<BM, BN, true>), grid, block, 0, st, A, B, C, m, n, k, lda,
                           ldb, ldc);
    else
        hipLaunchKernelGGL((dgemm_mfma<BM, BN, false>), grid, block, 0, st, A, B, C, m, n, k, lda,
                           ldb, ldc);
    return mgx::launch_status("dgemm_mfma");
}

int launch_dgemm(int tile, bool accumulate, const double *A, const double *B, double *C,
                 int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                 hipStream_t st) {
    if (tile == MGX_TILE_DEFAULT) tile = MGX_TILE_LARGE;
    switch (tile) {
        case MGX_TILE_LARGE:
            return launch_tile<128, 128>(accumulate, A, B, C, m, n, k, lda, ldb, ldc, st);
        case MGX_TILE_SMALL:
            return launch_tile<64, 64>(accumulate, A, B, C, m, n, k, lda, ldb, ldc, st);
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
    return launch_dgemm(tile, false, A, B, C, m, n, k, lda, ldb, ldc, (hipStream_t)stream);
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
    return launch_dgemm(tile, true, k_eff > 0 ? A + k_offset : A, B_chunk, C, m, n, k_eff, lda,
                        ldb, ldc, (hipStream_t)stream);
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
