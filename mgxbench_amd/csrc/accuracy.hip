// Accuracy report of a computed product against a double-double reference
// evaluated on the GPU for a sample of rows.
//
// Each lane owns one (row, column) of the sample and runs Dot2 (Ogita, Rump,
// Oishi 2005): TwoProduct by fma, Knuth's TwoSum, a compensated running sum,
// over k = 0..K-1 in order, plus sum |a||b| for the componentwise bound.
// Per-workgroup partial sums of the squared error, squared reference and the
// maxima go to library scratch and are folded on the host.
#include "mgx_internal.hpp"

#include <cmath>
#include <vector>

namespace {

constexpr int ER_THREADS = 256;

struct ErrPartial {
    double err2;      // sum (C - ref)^2
    double ref2;      // sum ref^2
    double comp;      // max |C - ref| / (|A||B|)
    double max_abs;   // max |C - ref|
};

__global__ __launch_bounds__(ER_THREADS) void dd_error_kernel(
    const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ C,
    int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc, const int64_t *__restrict__ rows,
    int64_t nrows, ErrPartial *partials) {
    __shared__ double sh[4][ER_THREADS / 64];
    const int64_t blocks_per_row = (n + ER_THREADS - 1) / ER_THREADS;
    const int64_t r = blockIdx.x / blocks_per_row;
    const int64_t col = (blockIdx.x % blocks_per_row) * ER_THREADS + threadIdx.x;
    double e2 = 0.0, f2 = 0.0, comp = 0.0, mabs = 0.0;
    if (r < nrows && col < n) {
        const int64_t row = rows[r];
        const double *a = A + row * lda;
        double p = 0.0, s = 0.0, absum = 0.0;
        for (int64_t kk = 0; kk < k; ++kk) {
            const double av = a[kk];
            const double bv = B[kk * ldb + col];
            const double h = av * bv;
            const double e = fma(av, bv, -h);
            const double x = p + h;
            const double z = x - p;
            const double q = (p - (x - z)) + (h - z);
            p = x;
            s = s + (q + e);
            absum = fma(fabs(av), fabs(bv), absum);
        }
        const double hi = p + s;
        const double z = hi - p;
        const double lo = (p - (hi - z)) + (s - z);
        const double c = C[row * ldc + col];
        const double d = (c - hi) - lo;
        e2 = d * d;
        f2 = hi * hi;
        mabs = fabs(d);
        comp = absum > 0.0 ? mabs / absum : (mabs > 0.0 ? INFINITY : 0.0);
        if (!(d == d)) {   // NaN result
            e2 = INFINITY;
            mabs = INFINITY;
            comp = INFINITY;
        }
    }
    double v[4] = {e2, f2, comp, mabs};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off);
            x = q < 2 ? x + y : fmax(x, y);
        }
        if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ErrPartial out{0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < ER_THREADS / 64; ++w) {
            out.err2 += sh[0][w];
            out.ref2 += sh[1][w];
            out.comp = fmax(out.comp, sh[2][w]);
            out.max_abs = fmax(out.max_abs, sh[3][w]);
        }
        partials[blockIdx.x] = out;
    }
}

}  // namespace

extern "C" int mgx_error_report(const double *A, const double *B, const double *C, int64_t m,
                                int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                                const int64_t *rows, int64_t nrows, mgx_error_stats *out) {
    mgx::clear_error();
    MGX_ARG_CHECK(out != nullptr, "out is NULL");
    MGX_ARG_CHECK(m > 0 && n > 0 && k >= 0, "invalid dimensions");
    MGX_ARG_CHECK(A && B && C && rows, "NULL pointer");
    MGX_ARG_CHECK(nrows > 0, "no rows to check");
    MGX_ARG_CHECK(lda >= k && ldb >= n && ldc >= n, "leading dimension too small");
    for (int64_t i = 0; i < nrows; ++i)
        MGX_ARG_CHECK(rows[i] >= 0 && rows[i] < m, "row index out of range");
    const int64_t blocks_per_row = (n + ER_THREADS - 1) / ER_THREADS;
    const int64_t nblocks = blocks_per_row * nrows;
    MGX_ARG_CHECK(nblocks <= 0x7fffffffll, "sample too large");
    const size_t rows_bytes = ((sizeof(int64_t) * nrows + 255) / 256) * 256;
    int st = MGX_OK;
    char *scratch = (char *)mgx::device_scratch(
        mgx::SCRATCH_VALIDATE, rows_bytes + sizeof(ErrPartial) * nblocks, nullptr, &st);
    if (!scratch) return st;
    int64_t *d_rows = (int64_t *)scratch;
    ErrPartial *d_part = (ErrPartial *)(scratch + rows_bytes);
    MGX_HIP_TRY(hipMemcpy(d_rows, rows, sizeof(int64_t) * nrows, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(dd_error_kernel, dim3((unsigned)nblocks), dim3(ER_THREADS), 0, 0, A, B, C,
                       n, k, lda, ldb, ldc, d_rows, nrows, d_part);
    if ((st = mgx::launch_status("dd_error_kernel")) != MGX_OK) return st;
    std::vector<ErrPartial> host(nblocks);
    MGX_HIP_TRY(hipMemcpy(host.data(), d_part, sizeof(ErrPartial) * nblocks,
                          hipMemcpyDeviceToHost));
    if ((st = mgx::scratch_release(mgx::SCRATCH_VALIDATE, nullptr)) != MGX_OK) return st;
    double e2 = 0.0, f2 = 0.0, comp = 0.0, mabs = 0.0;
    for (const auto &p : host) {
        e2 += p.err2;
        f2 += p.ref2;
        comp = std::fmax(comp, p.comp);
        mabs = std::fmax(mabs, p.max_abs);
    }
    out->normwise = f2 > 0.0 ? std::sqrt(e2 / f2) : std::sqrt(e2);
    out->componentwise = comp;
    out->max_abs = mabs;
    out->rows_checked = nrows;
    return MGX_OK;
}
