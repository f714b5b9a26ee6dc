// SUMMA panel streaming: the single-GPU stand-in for the reference's per-panel
// row/column broadcasts (2D-block-matrix/mmb.cu:239-361).
//
// One rank's local update C_ij += sum_p A_ik(p) . B_kj(p) over panels of
// width `panel` (ragged last panel, 2D-block-matrix/mmb.cu:242-246).  Panel
// p of A is an m x kc column block of the host A strip, panel p of B is kc
// rows of the host B strip.  They reach the device by hipMemcpy2DAsync /
// hipMemcpyAsync on a transfer stream; the compute stream waits on a
// per-buffer "panel ready" event and runs the accumulate product; a
// "compute done" event guards reuse of each of the two device buffers.
//
// Consecutive panels are gathered into a group buffer (whole panels covering
// mgx_split_group_k columns) and each group is one launch: FP64 runs the
// panel-fused kernel (mgx_dgemm_panels: one partial per panel added into the
// register-held C tile, so C is read and written once per group); split
// forms run one split product per group, so the exponent scan, operand
// conversion and (modular form) reconstruction of the whole C are paid once
// per group.  The copies overlap the previous group's product.  With pinned
// strips and asynchronous copies a group arrives as one A copy and one B
// copy; otherwise panel by panel, the reference's granularity.
//
// MI355X-native differences from the reference's loop:
//  * With pinned host strips the A panel is gathered by the DMA engine
//    (strided 2D copy) instead of host memcpy packing; with pageable strips
//    the library packs into its own pinned double buffer as the reference
//    does (pack_a / pack_b, 2D-block-matrix/mmb.cu:60-72).
//  * Buffer reuse is ordered by stream waits on events, not host syncs, so
//    the whole loop can be captured into one hipGraph and launched once.
#include "mgx_internal.hpp"

#include <chrono>
#include <cstring>
#include <vector>

extern "C" int mgx_split_gemm(int form, int slices, const double *A, const double *B, double *C,
                              int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                              int64_t ldc, int accumulate, void *workspace,
                              size_t workspace_bytes, void *stream);

namespace {

struct SummaResources {
    hipStream_t compute = nullptr, transfer = nullptr;
    hipEvent_t ready[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
    hipEvent_t staged[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join = nullptr;
    double *dA[2] = {nullptr, nullptr}, *dB[2] = {nullptr, nullptr};
    double *hA[2] = {nullptr, nullptr}, *hB[2] = {nullptr, nullptr};
    void *ws = nullptr;
    size_t ws_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~SummaResources() {
        if (compute) (void)hipStreamSynchronize(compute);
        if (transfer) (void)hipStreamSynchronize(transfer);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (int i = 0; i < 2; ++i) {
            if (ready[i]) (void)hipEventDestroy(ready[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (staged[i]) (void)hipEventDestroy(staged[i]);
            if (dA[i]) (void)hipFree(dA[i]);
            if (dB[i]) (void)hipFree(dB[i]);
            if (hA[i]) (void)hipHostFree(hA[i]);
            if (hB[i]) (void)hipHostFree(hB[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (ws) (void)hipFree(ws);
        if (compute) (void)hipStreamDestroy(compute);
        if (transfer) (void)hipStreamDestroy(transfer);
    }
};

bool host_is_pinned(const void *p) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeHost;
}

}  // namespace

extern "C" int64_t mgx_split_group_k(int form, int64_t panel, int64_t split_k, int64_t k) {
    if (panel <= 0 || k <= 0) return -1;
    if (form != MGX_FORM_FP64 && form != MGX_FORM_INT8 && form != MGX_FORM_INT8_MOD &&
        form != MGX_FORM_BF16)
        return -1;
    const int64_t p = panel < k ? panel : k;
    const int64_t want =
        split_k > 0 ? split_k : (form == MGX_FORM_FP64 ? MGX_FP64_GROUP_K : MGX_SPLIT_GROUP_K);
    const int64_t panels = (want + p - 1) / p;
    return panels * p < k ? panels * p : k;
}

extern "C" int mgx_summa_run(const mgx_summa_params *P, const double *hA, int64_t lda,
                             const double *hB, int64_t ldb, double *dC, int64_t ldc,
                             mgx_summa_stats *stats) {
    mgx::clear_error();
    MGX_ARG_CHECK(P != nullptr, "params is NULL");
    const int64_t m = P->m, n = P->n, k = P->k, bk = P->panel;
    MGX_ARG_CHECK(m > 0 && n > 0 && k > 0 && bk > 0, "sizes and panel width must be positive");
    MGX_ARG_CHECK(hA && hB && dC, "NULL matrix pointer");
    MGX_ARG_CHECK(lda >= k && ldb >= n && ldc >= n, "leading dimension too small");
    MGX_ARG_CHECK(P->form == MGX_FORM_FP64 || P->form == MGX_FORM_INT8 ||
                      P->form == MGX_FORM_INT8_MOD || P->form == MGX_FORM_BF16,
                  "unsupported form");
    MGX_ARG_CHECK(P->copy_mode == MGX_SUMMA_COPY_ASYNC || P->copy_mode == MGX_SUMMA_COPY_HOST_SYNC,
                  "unknown copy mode");
    const int64_t panel = bk < k ? bk : k;
    const int64_t npanels = (k + panel - 1) / panel;
    // group width: one panel (FP64) or whole panels per split product
    const int64_t gw = mgx_split_group_k(P->form, panel, P->split_k, k);
    const int64_t ngroups = (k + gw - 1) / gw;
    const bool pinned = host_is_pinned(hA) && host_is_pinned(hB);
    const bool graph = P->use_graph != 0;
    MGX_ARG_CHECK(!graph || (pinned && P->copy_mode == MGX_SUMMA_COPY_ASYNC && !P->serialize),
                  "hipGraph capture needs pinned host strips, async copies and no serialization");

    SummaResources R;
    MGX_HIP_TRY(hipStreamCreateWithFlags(&R.compute, hipStreamNonBlocking));
    MGX_HIP_TRY(hipStreamCreateWithFlags(&R.transfer, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        MGX_HIP_TRY(hipEventCreateWithFlags(&R.ready[i], hipEventDisableTiming));
        MGX_HIP_TRY(hipEventCreateWithFlags(&R.done[i], hipEventDisableTiming));
        MGX_HIP_TRY(hipEventCreateWithFlags(&R.staged[i], hipEventDisableTiming));
        if (hipMalloc(&R.dA[i], sizeof(double) * m * gw) != hipSuccess ||
            hipMalloc(&R.dB[i], sizeof(double) * gw * n) != hipSuccess)
            return mgx::set_error(MGX_ERR_ALLOC, "panel buffer allocation failed");
        if (!pinned) {
            if (hipHostMalloc((void **)&R.hA[i], sizeof(double) * m * panel) != hipSuccess ||
                hipHostMalloc((void **)&R.hB[i], sizeof(double) * panel * n) != hipSuccess)
                return mgx::set_error(MGX_ERR_ALLOC, "pinned staging allocation failed");
        }
    }
    MGX_HIP_TRY(hipEventCreateWithFlags(&R.fork, hipEventDisableTiming));
    MGX_HIP_TRY(hipEventCreateWithFlags(&R.join, hipEventDisableTiming));
    if (P->form != MGX_FORM_FP64) {
        R.ws_bytes = mgx_split_workspace_bytes(P->form, P->slices, m, n, gw);
        if (R.ws_bytes == 0) return mgx::set_error(MGX_ERR_UNSUPPORTED, "slice count out of range");
        if (hipMalloc(&R.ws, R.ws_bytes) != hipSuccess)
            return mgx::set_error(MGX_ERR_ALLOC, "split workspace allocation failed");
    }

    // Enqueues the panel loop on the two streams.  Device group buffers are
    // reused under stream waits on "compute done"; with pageable strips the
    // host packs each panel into one of two pinned staging buffers after the
    // copy that last read it has finished.
    auto enqueue = [&]() -> int {
        int rc;
        MGX_HIP_TRY(hipEventRecord(R.fork, R.compute));
        MGX_HIP_TRY(hipStreamWaitEvent(R.transfer, R.fork, 0));
        int64_t pi = 0;   // panel counter (staging buffer parity)
        for (int64_t g = 0; g < ngroups; ++g) {
            const int cur = (int)(g & 1);
            const int64_t g0 = g * gw;
            const int64_t gk = (k - g0) < gw ? (k - g0) : gw;
            if (g >= 2) MGX_HIP_TRY(hipStreamWaitEvent(R.transfer, R.done[cur], 0));
            if (pinned && gk > panel && P->copy_mode == MGX_SUMMA_COPY_ASYNC && !P->serialize) {
                // a split group's panels are adjacent columns of the A strip
                // and adjacent rows of the B strip: one copy each (a copy
                // costs ~50 us of idle copy engine whatever its size)
                MGX_HIP_TRY(hipMemcpy2DAsync(R.dA[cur], sizeof(double) * gk, hA + g0,
                                             sizeof(double) * lda, sizeof(double) * gk, m,
                                             hipMemcpyHostToDevice, R.transfer));
                MGX_HIP_TRY(hipMemcpy2DAsync(R.dB[cur], sizeof(double) * n, hB + g0 * ldb,
                                             sizeof(double) * ldb, sizeof(double) * n, gk,
                                             hipMemcpyHostToDevice, R.transfer));
                pi += (gk + panel - 1) / panel;
            } else {
                for (int64_t j0 = 0; j0 < gk; j0 += panel, ++pi) {
                    const int64_t kc = (gk - j0) < panel ? (gk - j0) : panel;
                    const int64_t k0 = g0 + j0;
                    const double *srcA = hA + k0;
                    const double *srcB = hB + k0 * ldb;
                    int64_t spitchA = lda, spitchB = ldb;
                    const int s = (int)(pi & 1);
                    if (!pinned) {
                        // pack_a / pack_b into the pinned staging buffer
                        if (pi >= 2) MGX_HIP_TRY(hipEventSynchronize(R.staged[s]));
                        for (int64_t i = 0; i < m; ++i)
                            std::memcpy(R.hA[s] + i * kc, hA + i * lda + k0, sizeof(double) * kc);
                        for (int64_t r = 0; r < kc; ++r)
                            std::memcpy(R.hB[s] + r * n, hB + (k0 + r) * ldb, sizeof(double) * n);
                        srcA = R.hA[s];
                        srcB = R.hB[s];
                        spitchA = kc;
                        spitchB = n;
                    }
                    // panel j0 of the group: columns j0.. of the m x gk A buffer
                    // (lda = gk) and rows j0.. of the gk x n B buffer
                    MGX_HIP_TRY(hipMemcpy2DAsync(R.dA[cur] + j0, sizeof(double) * gk, srcA,
                                                 sizeof(double) * spitchA, sizeof(double) * kc, m,
                                                 hipMemcpyHostToDevice, R.transfer));
                    MGX_HIP_TRY(hipMemcpy2DAsync(R.dB[cur] + j0 * n, sizeof(double) * n, srcB,
                                                 sizeof(double) * spitchB, sizeof(double) * n, kc,
                                                 hipMemcpyHostToDevice, R.transfer));
                    if (!pinned) MGX_HIP_TRY(hipEventRecord(R.staged[s], R.transfer));
                    if (P->copy_mode == MGX_SUMMA_COPY_HOST_SYNC || P->serialize)
                        MGX_HIP_TRY(hipStreamSynchronize(R.transfer));
                }
            }
            MGX_HIP_TRY(hipEventRecord(R.ready[cur], R.transfer));
            MGX_HIP_TRY(hipStreamWaitEvent(R.compute, R.ready[cur], 0));
            if (P->form == MGX_FORM_FP64) {
                // one launch per group, one partial per panel inside it
                rc = mgx_dgemm_panels(MGX_TILE_DEFAULT, 1, R.dA[cur], R.dB[cur], dC, m, n, gk,
                                      panel, gk, n, ldc, R.compute);
            } else {
                rc = mgx_split_gemm(P->form, P->slices, R.dA[cur], R.dB[cur], dC, m, n, gk,
                                    gk, n, ldc, 1, R.ws, R.ws_bytes, R.compute);
            }
            if (rc != MGX_OK) return rc;
            MGX_HIP_TRY(hipEventRecord(R.done[cur], R.compute));
            if (P->serialize) MGX_HIP_TRY(hipStreamSynchronize(R.compute));
        }
        MGX_HIP_TRY(hipEventRecord(R.join, R.transfer));
        MGX_HIP_TRY(hipStreamWaitEvent(R.compute, R.join, 0));
        return MGX_OK;
    };

    int rc;
    if (graph) {
        MGX_HIP_TRY(hipStreamBeginCapture(R.compute, hipStreamCaptureModeThreadLocal));
        rc = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(R.compute, &g);
        R.graph = g;
        if (rc != MGX_OK) return rc;
        if (e != hipSuccess)
            return mgx::set_error(MGX_ERR_HIP, std::string("hipStreamEndCapture failed: ") +
                                                   hipGetErrorString(e));
        MGX_HIP_TRY(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
        MGX_HIP_TRY(hipGraphUpload(R.exec, R.compute));
    }
    MGX_HIP_TRY(hipDeviceSynchronize());
    const auto t0 = std::chrono::steady_clock::now();
    if (graph) {
        MGX_HIP_TRY(hipGraphLaunch(R.exec, R.compute));
    } else {
        if ((rc = enqueue()) != MGX_OK) return rc;
    }
    MGX_HIP_TRY(hipStreamSynchronize(R.compute));
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
        stats->panels = npanels;
        stats->host_pinned = pinned ? 1 : 0;
    }
    return MGX_OK;
}

extern "C" int mgx_host_alloc(void **ptr, size_t bytes) {
    mgx::clear_error();
    MGX_ARG_CHECK(ptr != nullptr, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return MGX_OK;
    if (hipHostMalloc(ptr, bytes) != hipSuccess) {
        *ptr = nullptr;
        return mgx::set_error(MGX_ERR_ALLOC, "hipHostMalloc failed");
    }
    return MGX_OK;
}

extern "C" int mgx_host_free(void *ptr) {
    mgx::clear_error();
    if (ptr) MGX_HIP_TRY(hipHostFree(ptr));
    return MGX_OK;
}
