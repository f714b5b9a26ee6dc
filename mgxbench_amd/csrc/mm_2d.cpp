// mm_2d: the 2D SUMMA executable of MGXBench for one rank on one MI355X.
//
// Keeps the reference CLI and result contract (2D-block-matrix/mmb.cu:74-101,
// 377-401):  mm_2d <auto|device> <N> <MMM|YYY|CCC|XXX|NNN|WWW|SSS>
// env MGXBENCH_GRID_ROWS / MGXBENCH_GRID_COLS / MGXBENCH_SUMMA_BK; one
// BENCHMARK_RESULT line; exit 0 when valid, 1 otherwise.
//
// The rank's local update C_ij += sum_p A_ik(p) . B_kj(p) runs through
// mgx_summa_run: panels stream from the host strips on a transfer stream,
// overlapped with the accumulate product on a compute stream (the single-GPU
// stand-in for the per-panel broadcasts).  As in the reference's timed loop
// (2D-block-matrix/mmb.cu:264-265), the strips are pageable and every panel
// is packed on the host into a pinned staging buffer before its copy (here
// by a pool of host threads, MGX_HOST_THREADS, default min(16, cores)).  With
// one rank, every library code moves panels host -> device; MMM/YYY/NNN
// overlap the copies, CCC/XXX/WWW/SSS wait for each copy on the host, as the
// reference's P=1 paths do.
//
// Extensions (new env vars; with none of them set the result line is the
// reference's byte for byte, otherwise it carries the extra keys form=,
// slices=, share= and panels=): MGX_FORM=fp64|int8|bf16|mod and MGX_SLICES
// select the product form; MGX_INIT=uniform:SEED replaces the constant A=1,
// B=2 inputs (validated by the GPU double-double check instead of the
// constant rule); MGX_SHARE_GRID=RxC computes rank (0,0)'s share of an R x C
// grid at this N (m = N/R, n = N/C, K = N); MGX_SUMMA_GRAPH=1 captures the
// panel loop in a hipGraph (with pinned strips); MGX_HOST_PINNED=1 allocates
// the strips pinned, so panels are gathered by strided DMA without host
// packing.
// MGX_TEST_PERTURB=<index>:<value> (test hook) overwrites one element of the
// local C before validation.
#include <cmath>
#include <vector>

#include "drivers_common.hpp"

using namespace mgxdrv;

static void fail(const char *msg) {
    std::fprintf(stderr, "2D SUMMA: %s\n", msg);
    std::exit(2);
}

static void grid_shape(int p, int *pr, int *pc) {
    const int r = env_int("MGXBENCH_GRID_ROWS", 0);
    const int c = env_int("MGXBENCH_GRID_COLS", 0);
    if (r < 0 || c < 0) { *pr = 0; *pc = 0; return; }
    if (r && c) { *pr = r; *pc = c; return; }
    if (r) { *pr = r; *pc = (p % r == 0 ? p / r : 0); return; }
    if (c) { *pc = c; *pr = (p % c == 0 ? p / c : 0); return; }
    // near-square factorisation, rows <= cols
    int a = (int)std::sqrt((double)p);
    while (a > 1 && p % a) --a;
    *pr = a;
    *pc = p / a;
}

int main(int argc, char **argv) {
    if (argc != 4) fail("Usage: mm_2d <auto|device> <N> <MMM|YYY|CCC|XXX|NNN|WWW|SSS>");
    const int N = std::atoi(argv[2]);
    const char *lib = argv[3];
    const char *codes[] = {"MMM", "YYY", "CCC", "XXX", "NNN", "WWW", "SSS"};
    bool supported = false;
    for (const char *c : codes) supported |= std::strcmp(lib, c) == 0;
    if (!supported || N <= 0) fail("invalid matrix size or library");

    const int p = launcher_world_size();
    if (p != 1) fail("this build runs one rank per GPU process; the multi-rank exchange is not built");
    int Pr, Pc;
    grid_shape(p, &Pr, &Pc);
    if (Pr <= 0 || Pc <= 0 || (long long)Pr * Pc != p)
        fail("invalid process grid: rows*cols must equal MPI ranks");

    // rank (0,0)'s share of an emulated R x C grid
    int Sr = 1, Sc = 1;
    const std::string share = env_str("MGX_SHARE_GRID", "1x1");
    if (std::sscanf(share.c_str(), "%dx%d", &Sr, &Sc) != 2 || Sr <= 0 || Sc <= 0)
        fail("MGX_SHARE_GRID must be RxC");
    if (N % (Pr * Sr) || N % (Pc * Sc)) fail("N must be divisible by both process-grid dimensions");
    const int64_t m = N / (Pr * Sr);
    const int64_t n = N / (Pc * Sc);
    int BK = env_int("MGXBENCH_SUMMA_BK", 256);
    if (BK <= 0) fail("MGXBENCH_SUMMA_BK must be positive");
    if (BK > N) BK = N;

    Options opt;
    if (!parse_options(&opt)) fail("invalid MGX_FORM or MGX_INIT");
    int dev = 0;
    if (select_device(argv[1], &dev) != 0) fail("GPU device mapping invalid");
    const bool graph = env_int("MGX_SUMMA_GRAPH", 0) != 0;
    // a captured panel loop gathers its panels by DMA from pinned strips
    const bool pinned = env_int("MGX_HOST_PINNED", graph ? 1 : 0) != 0;

    std::printf("SUMMA rectangular grid=%dx%d N=%d BK=%d library=%s ranks=%d pipeline=2\n", Pr, Pc, N,
                BK, lib, p);
    const bool extras = extras_requested();
    if (extras)
        std::printf("form=%s slices=%d share=%dx%d local C %lldx%lld host strips %s%s\n",
                    form_name(opt.form), opt.slices, Sr, Sc, (long long)m, (long long)n,
                    pinned ? "pinned" : "pageable", graph ? " hipGraph" : "");

    // Host strips: A_ik (m x N) and B_kj (N x n) of the rank's row / column.
    const size_t a_elems = (size_t)m * N, b_elems = (size_t)N * n;
    double *hA = nullptr, *hB = nullptr;
    std::vector<double> pageA, pageB;
    if (pinned) {
        if (mgx_host_alloc((void **)&hA, a_elems * 8) != MGX_OK ||
            mgx_host_alloc((void **)&hB, b_elems * 8) != MGX_OK)
            fail("pinned host allocation failed");
    } else {
        pageA.resize(a_elems);
        pageB.resize(b_elems);
        hA = pageA.data();
        hB = pageB.data();
    }
    for (int64_t i = 0; i < m; ++i)
        for (int64_t kk = 0; kk < N; ++kk)
            hA[i * N + kk] = opt.uniform ? uniform_at(opt.seed, (uint64_t)(i * N + kk)) : 1.0;
    for (int64_t kk = 0; kk < N; ++kk)
        for (int64_t j = 0; j < n; ++j)
            hB[kk * n + j] = opt.uniform ? uniform_at(opt.seed + 1, (uint64_t)(kk * N + j)) : 2.0;

    double *dC = nullptr;
    if (hipMalloc((void **)&dC, sizeof(double) * m * n) != hipSuccess) fail("hipMalloc failed");
    if (mgx_fill(dC, m, n, n, 0, 0, 0.0, nullptr) != MGX_OK) fail(mgx_last_error());

    mgx_summa_params P{};
    P.m = m;
    P.n = n;
    P.k = N;
    P.panel = BK;
    P.form = opt.form;
    P.slices = opt.slices;
    const bool async_copy = !std::strcmp(lib, "MMM") || !std::strcmp(lib, "YYY") ||
                            !std::strcmp(lib, "NNN");
    P.copy_mode = async_copy ? MGX_SUMMA_COPY_ASYNC : MGX_SUMMA_COPY_HOST_SYNC;
    P.use_graph = graph ? 1 : 0;
    P.serialize = env_int("MGX_SUMMA_SERIALIZE", 0) != 0;
    P.split_k = env_int("MGX_SPLIT_K", 0);
    P.host_threads = env_int("MGX_HOST_THREADS", 0);
    mgx_summa_stats S;
    if (mgx_summa_run(&P, hA, N, hB, n, dC, n, &S) != MGX_OK) fail(mgx_last_error());
    const double elapsed = S.seconds;
    if (!apply_test_perturbation(dC, m * n)) fail("invalid MGX_TEST_PERTURB");

    int valid = 1;
    double normwise = -1.0;
    if (opt.uniform) {
        double *dA = nullptr, *dB = nullptr;
        if (hipMalloc((void **)&dA, a_elems * 8) != hipSuccess ||
            hipMalloc((void **)&dB, b_elems * 8) != hipSuccess ||
            hipMemcpy(dA, hA, a_elems * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dB, hB, b_elems * 8, hipMemcpyHostToDevice) != hipSuccess)
            fail("validation upload failed");
        valid = check_uniform(dA, dB, dC, m, n, N, N, n, n, &normwise);
        (void)hipFree(dA);
        (void)hipFree(dB);
    } else {
        // 2D host rule: every element finite and |C - 2N| <= 1e-9 + 1e-12 |2N|
        std::vector<double> localC((size_t)m * n);
        if (hipMemcpy(localC.data(), dC, localC.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
            fail("hipMemcpy failed");
        const double expected = 2.0 * N;
        const double tolerance = 1e-9 + 1e-12 * std::fabs(expected);
        for (size_t i = 0; i < localC.size(); ++i) {
            const double v = localC[i];
            if (!std::isfinite(v) || std::fabs(v - expected) > tolerance) {
                std::fprintf(stderr,
                             "Rank 0: 2D validation failed at local index=%zu: got %.17g expected "
                             "%.17g\n",
                             i, v, expected);
                valid = 0;
                break;
            }
        }
    }
    std::printf("BENCHMARK_RESULT decomposition=2D library=%s size=%d ranks=%d seconds=%.9f valid=%d "
                "loops=1 grid=%dx%d",
                lib, N, p, elapsed, valid, Pr, Pc);
    if (extras)
        std::printf(" form=%s slices=%d share=%dx%d panels=%lld", form_name(opt.form), opt.slices,
                    Sr, Sc, (long long)S.panels);
    std::printf("\n");
    std::fflush(stdout);
    (void)hipFree(dC);
    if (pinned) {
        mgx_host_free(hA);
        mgx_host_free(hB);
    }
    return valid ? 0 : 1;
}
