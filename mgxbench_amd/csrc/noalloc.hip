// The no-allocation split product (mgx_split_gemm_noalloc).
//
// Serves callers that cannot allocate device memory (code under stream
// capture, tight serving loops): its workspace is this file's static device
// array, MGX_NOALLOC_SCRATCH_BYTES per GPU, held from the loading of this
// file's code object (on the first call) until the process ends.  It lives in
// its own source file so that only callers of this entry point load it.
// Host code never passes the array's name to a kernel: hipGetSymbolAddress
// gives each device's own address once, with that device current, and that
// address is the workspace of the split product.  The split product writes
// every byte of the workspace it reads (operand planes with their padding,
// residues, exponents) and clears the row/column maxima with a fill kernel,
// so no contents carry over between calls; calls that overlap on one GPU
// share the array and race.
#include "mgx_internal.hpp"

#include <mutex>
#include <string>

static __device__ __align__(256) char scratch[MGX_NOALLOC_SCRATCH_BYTES];

namespace mgx {
int split_gemm_run(int form, int slices, int tile, const double *A, const double *B, double *C,
                   int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                   int accumulate, void *workspace, size_t workspace_bytes, void *stream,
                   void *const *events);
}

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(void **out) {
    int dev = 0;
    MGX_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return mgx::set_error(MGX_ERR_UNSUPPORTED, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return mgx::set_error(MGX_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                                   hipGetErrorString(e));
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return MGX_OK;
}
}  // namespace

extern "C" int mgx_split_gemm_noalloc(int form, int slices, const double *A, const double *B,
                                      double *C, int64_t m, int64_t n, int64_t k, int64_t lda,
                                      int64_t ldb, int64_t ldc, int accumulate, void *stream) {
    mgx::clear_error();
    const size_t need = mgx_split_workspace_bytes(form, slices, m, n, k);
    if (need > (size_t)MGX_NOALLOC_SCRATCH_BYTES)
        return mgx::set_error(MGX_ERR_WORKSPACE,
                              "split product of this size needs " + std::to_string(need) +
                                  " bytes of workspace, beyond the no-allocation scratch of " +
                                  std::to_string((size_t)MGX_NOALLOC_SCRATCH_BYTES) + " bytes");
    void *ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc != MGX_OK) return rc;
    return mgx::split_gemm_run(form, slices, MGX_TILE_DEFAULT, A, B, C, m, n, k, lda, ldb, ldc,
                               accumulate, ws, (size_t)MGX_NOALLOC_SCRATCH_BYTES, stream, nullptr);
}
