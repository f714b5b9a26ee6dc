// Autotuner measurement: time one candidate configuration of the local
// product (form, slices, tile, panel width, split group) on device-resident
// operands.
#include "mgx_internal.hpp"

namespace {

int run_once(const mgx_tune_config *c, const double *A, const double *B, double *C, int64_t m,
             int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc, void *ws,
             size_t ws_bytes, hipStream_t st) {
    // whole-panel groups, one launch each (the SUMMA loop's compute side):
    // FP64 groups run the panel-fused kernel, one partial per panel
    const int64_t group = mgx_split_group_k(c->form, c->panel, c->split_k, k);
    for (int64_t k0 = 0; k0 < k; k0 += group) {
        const int64_t kc = (k - k0) < group ? (k - k0) : group;
        const int acc = k0 > 0 ? 1 : 0;
        int rc;
        if (c->form == MGX_FORM_FP64)
            rc = mgx_dgemm_panels(c->tile, acc, A + k0, B + k0 * ldb, C, m, n, kc, c->panel, lda,
                                  ldb, ldc, st);
        else
            rc = mgx_split_gemm_ex(c->form, c->slices, c->tile, A + k0, B + k0 * ldb, C, m, n, kc,
                                   lda, ldb, ldc, acc, ws, ws_bytes, st);
        if (rc != MGX_OK) return rc;
    }
    return MGX_OK;
}

}  // namespace

extern "C" int mgx_time_local_product(const mgx_tune_config *cfg, const double *A,
                                      const double *B, double *C, int64_t m, int64_t n, int64_t k,
                                      int64_t lda, int64_t ldb, int64_t ldc, void *workspace,
                                      size_t workspace_bytes, int warmup, int reps,
                                      double *ms_out) {
    mgx::clear_error();
    MGX_ARG_CHECK(cfg && ms_out, "NULL pointer");
    MGX_ARG_CHECK(cfg->form == MGX_FORM_FP64 || cfg->form == MGX_FORM_INT8 ||
                      cfg->form == MGX_FORM_INT8_MOD || cfg->form == MGX_FORM_BF16,
                  "unsupported form");
    MGX_ARG_CHECK(cfg->panel > 0 && m > 0 && n > 0 && k > 0, "invalid sizes");
    MGX_ARG_CHECK(warmup >= 0 && reps > 0, "invalid repetition counts");
    hipStream_t st = nullptr;
    MGX_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MGX_OK;
    auto cleanup = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(st);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return mgx::set_error(MGX_ERR_HIP, "hipEventCreate failed");
    }
    for (int i = 0; i < warmup && rc == MGX_OK; ++i)
        rc = run_once(cfg, A, B, C, m, n, k, lda, ldb, ldc, workspace, workspace_bytes, st);
    if (rc == MGX_OK && hipEventRecord(e0, st) != hipSuccess) rc = MGX_ERR_HIP;
    for (int i = 0; i < reps && rc == MGX_OK; ++i)
        rc = run_once(cfg, A, B, C, m, n, k, lda, ldb, ldc, workspace, workspace_bytes, st);
    if (rc == MGX_OK && hipEventRecord(e1, st) != hipSuccess) rc = MGX_ERR_HIP;
    float ms = 0.0f;
    if (rc == MGX_OK && (hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = mgx::set_error(MGX_ERR_HIP, "event timing failed");
    cleanup();
    if (rc != MGX_OK) return rc;
    *ms_out = (double)ms / reps;
    return MGX_OK;
}
