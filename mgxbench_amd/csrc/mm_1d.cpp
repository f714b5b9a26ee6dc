// mm_1d: the 1D row-block executable of MGXBench for one rank on one MI355X.
//
// Keeps the reference CLI and result contract (1D-row-block/mmb.cu:178-236,
// 927-967, 1036):  mm_1d <deviceId|auto> <N> <ABC>  with A, B, C each one of
// M Y C X N S W; env MGXBENCH_INNER_LOOPS (default 10), MGXBENCH_INTERNAL_WARMUP
// (default 0), MM_PIPELINE_K_CHUNK (default 2048); the validation block of
// validate_matrix_C; one BENCHMARK_RESULT line; exit 0 when valid, 2 otherwise.
//
// With one rank the A scatter and C gather phases vanish (the reference
// aliases the local and global buffers, 1D-row-block/mmb.cu:403-427) and the
// B phase reduces to (1D-row-block/mmb.cu:694-858):
//   M  copy B device -> host each iteration, then the full product;
//   C, N, W  the full product (the one-rank broadcast moves nothing);
//   S  K-chunk pipeline, chunk copies device -> device on a side stream;
//   Y  K-chunk pipeline, chunk device -> host -> device through pinned
//      buffers on a side stream;
//   X  K-chunk pipeline reading B in place.
// The pipeline double-buffers chunk stages with "B ready" / "compute done"
// events and runs C += A[:, k0:k0+kc] . B_chunk per chunk on a compute stream.
//
// Extensions (new env vars): MGX_FORM=fp64|int8|bf16|mod, MGX_SLICES, MGX_INIT=
// uniform:SEED (validated by the GPU double-double check), MGX_SHARE_RANKS=P
// (rank 0's row block of a P-rank run: A_i is N/P x N), MGX_SPLIT_K.  Only
// when one of them is set does the result line carry the extra keys form=,
// slices= and share_ranks=; otherwise it is the reference's line exactly.
// MGX_TEST_PERTURB=<index>:<value> (test hook) overwrites one element of C
// before validation.
#include <chrono>
#include <vector>

#include "drivers_common.hpp"

using namespace mgxdrv;

// Default columns per split product in the K-chunk pipeline.
constexpr int64_t PIPELINE_SPLIT_K = 4096;

#define TILE_DIM 32

static void fail(const char *msg, int code = 1) {
    std::fprintf(stderr, "mm_1d: %s\n", msg);
    std::exit(code);
}

#define DRV_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "[MGX rank=0] HIP failure %s:%d (%s): %s\n", __FILE__, \
                         __LINE__, #call, hipGetErrorString(e_));                      \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

#define DRV_MGX(call)                                                                  \
    do {                                                                               \
        if ((call) != MGX_OK) fail(mgx_last_error());                                  \
    } while (0)

static bool valid_library(char c) { return std::strchr("MYCXNSW", c) != nullptr && c != 0; }

static int pipeline_k_chunk(int k) {
    int chunk = 2048;
    const int requested = env_int("MM_PIPELINE_K_CHUNK", 0);
    if (requested > 0) chunk = requested;
    if (chunk > k) chunk = k;
    if (chunk >= TILE_DIM) chunk = (chunk / TILE_DIM) * TILE_DIM;
    return chunk > 0 ? chunk : k;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr,
                     "Usage: %s <deviceId> <matrix_size> <ABC libraries>\n"
                     "  M=MPI-SYNC, Y=MPI-ASYNC, C=CUDA-Aware MPI-SYNC, X=CUDA-Aware MPI-ASYNC, "
                     "N=NCCL-SYNC, S=NVSHMEM-ASYNC, W=NVSHMEM-SYNC\n"
                     "  Example: %s 0 8192 SNS\n",
                     argv[0], argv[0]);
        return 1;
    }
    const int nranks = launcher_world_size();
    if (nranks != 1) fail("this build runs one rank per GPU process; the multi-rank exchange is not built");
    const int matrix_size = std::atoi(argv[2]);
    char libs[4] = {0, 0, 0, 0};
    std::strncpy(libs, argv[3], 3);
    if (std::strlen(argv[3]) != 3 || !valid_library(libs[0]) || !valid_library(libs[1]) ||
        !valid_library(libs[2])) {
        std::fprintf(stderr, "Invalid communication string. Use exactly 3 characters from M, Y, C, X, N, S, W.\n");
        return 1;
    }
    if (matrix_size <= 0) fail("invalid matrix size");
    const char libB = libs[1];

    Options opt;
    if (!parse_options(&opt)) fail("invalid MGX_FORM or MGX_INIT");
    const int share = env_int("MGX_SHARE_RANKS", 1);
    if (share <= 0 || matrix_size % share) fail("MGX_SHARE_RANKS must divide N");
    int dev = 0;
    if (select_device(argv[1], &dev) != 0) fail("No unique visible GPU", 3);

    const int64_t k = matrix_size, n = matrix_size;
    const int64_t mi = matrix_size / share;
    const size_t bytes_A = sizeof(double) * mi * k, bytes_B = sizeof(double) * k * n,
                 bytes_C = sizeof(double) * mi * n;
    double *d_A = nullptr, *d_B = nullptr, *d_C = nullptr;
    DRV_HIP(hipMalloc((void **)&d_A, bytes_A));
    DRV_HIP(hipMalloc((void **)&d_B, bytes_B));
    DRV_HIP(hipMalloc((void **)&d_C, bytes_C));
    if (opt.uniform) {
        DRV_MGX(mgx_fill(d_A, mi, k, k, 1, opt.seed, 0.0, nullptr));
        DRV_MGX(mgx_fill(d_B, k, n, n, 1, opt.seed + 1, 0.0, nullptr));
    } else {
        DRV_MGX(mgx_fill(d_A, mi, k, k, 0, 0, 1.0, nullptr));
        DRV_MGX(mgx_fill(d_B, k, n, n, 0, 0, 2.0, nullptr));
    }
    DRV_HIP(hipDeviceSynchronize());

    const int loop_count = env_int("MGXBENCH_INNER_LOOPS", 10);
    if (loop_count < 1) fail("MGXBENCH_INNER_LOOPS must be positive", 4);
    const int warmup_count = env_int("MGXBENCH_INTERNAL_WARMUP", 0);
    if (warmup_count < 0) fail("MGXBENCH_INTERNAL_WARMUP must be non-negative", 4);

    const bool pipelined = libB == 'S' || libB == 'Y' || libB == 'X';
    // split forms stage whole groups of chunks per split product (MGX_SPLIT_K
    // columns, default PIPELINE_SPLIT_K): the chunk copies here are short
    // device-side copies, so the products set the pace and wide groups
    // amortise a split product's fixed costs (exponent scan, whole-C
    // reconstruction) over more depth
    const int kchunk = opt.form == MGX_FORM_FP64
                           ? pipeline_k_chunk((int)k)
                           : (int)mgx_split_group_k(opt.form, pipeline_k_chunk((int)k),
                                                    env_int("MGX_SPLIT_K", PIPELINE_SPLIT_K), k);
    hipStream_t comm = nullptr, compute = nullptr;
    hipEvent_t b_ready[2] = {}, compute_done[2] = {};
    double *b_stage[2] = {nullptr, nullptr};
    double *y_host[2] = {nullptr, nullptr};
    std::vector<double> host_B;
    if (pipelined) {
        DRV_HIP(hipStreamCreateWithFlags(&comm, hipStreamNonBlocking));
        DRV_HIP(hipStreamCreateWithFlags(&compute, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            DRV_HIP(hipEventCreateWithFlags(&b_ready[i], hipEventDisableTiming));
            DRV_HIP(hipEventCreateWithFlags(&compute_done[i], hipEventDisableTiming));
            if (libB != 'X') DRV_HIP(hipMalloc((void **)&b_stage[i], sizeof(double) * kchunk * n));
            if (libB == 'Y') DRV_MGX(mgx_host_alloc((void **)&y_host[i], sizeof(double) * kchunk * n));
        }
    } else if (libB == 'M') {
        host_B.resize((size_t)k * n);
    }
    void *ws = nullptr;
    size_t ws_bytes = 0;
    if (opt.form != MGX_FORM_FP64) {
        ws_bytes = mgx_split_workspace_bytes(opt.form, opt.slices, mi, n, pipelined ? kchunk : k);
        if (ws_bytes == 0) fail("MGX_SLICES out of range");
        DRV_HIP(hipMalloc(&ws, ws_bytes));
    }

    // C += A[:, k_off:k_off+kc] . Bc on `st` in the selected form.
    auto accumulate = [&](const double *Bc, int64_t k_off, int64_t kc, hipStream_t st) {
        if (opt.form == MGX_FORM_FP64)
            DRV_MGX(mgx_dgemm_accumulate(d_A, Bc, d_C, mi, n, k, k_off, kc, k, n, n, TILE_DIM, st));
        else
            DRV_MGX(mgx_split_gemm(opt.form, opt.slices, d_A + k_off, Bc, d_C, mi, n, kc, k, n,
                                   n, 1, ws, ws_bytes, st));
    };

    auto one_iteration = [&]() {
        if (!pipelined) {
            if (libB == 'M') DRV_HIP(hipMemcpy(host_B.data(), d_B, bytes_B, hipMemcpyDeviceToHost));
            if (opt.form == MGX_FORM_FP64)
                DRV_MGX(mgx_dgemm(d_A, d_B, d_C, mi, n, k, k, n, n, TILE_DIM, nullptr));
            else
                DRV_MGX(mgx_split_gemm(opt.form, opt.slices, d_A, d_B, d_C, mi, n, k, k, n, n, 0,
                                       ws, ws_bytes, nullptr));
            DRV_HIP(hipGetLastError());
            DRV_HIP(hipDeviceSynchronize());
            return;
        }
        DRV_MGX(mgx_fill(d_C, mi, n, n, 0, 0, 0.0, compute));
        const int num_chunks = (int)((k + kchunk - 1) / kchunk);
        bool used[2] = {false, false};
        // stage chunk c (offset off, depth kc) into buffer `buf` on the comm stream
        auto stage = [&](int buf, int64_t off, int64_t kc) {
            const size_t count = (size_t)kc * n;
            if (used[buf]) DRV_HIP(hipStreamWaitEvent(comm, compute_done[buf], 0));
            if (libB == 'S') {
                DRV_HIP(hipMemcpyAsync(b_stage[buf], d_B + off * n, count * 8, hipMemcpyDeviceToDevice,
                                       comm));
            } else {  // Y: device -> pinned host (waited), host -> device stage
                DRV_HIP(hipMemcpyAsync(y_host[buf], d_B + off * n, count * 8, hipMemcpyDeviceToHost,
                                       comm));
                DRV_HIP(hipStreamSynchronize(comm));
                DRV_HIP(hipMemcpyAsync(b_stage[buf], y_host[buf], count * 8, hipMemcpyHostToDevice,
                                       comm));
            }
            DRV_HIP(hipEventRecord(b_ready[buf], comm));
            used[buf] = true;
        };
        if (libB != 'X') stage(0, 0, kchunk < k ? kchunk : k);
        for (int c = 0; c < num_chunks; ++c) {
            const int cur = c & 1;
            const int64_t off = (int64_t)c * kchunk;
            const int64_t kc = (k - off) < kchunk ? (k - off) : kchunk;
            const double *Bc = d_B + off * n;
            if (libB != 'X') {
                DRV_HIP(hipStreamWaitEvent(compute, b_ready[cur], 0));
                Bc = b_stage[cur];
            }
            accumulate(Bc, off, kc, compute);
            DRV_HIP(hipEventRecord(compute_done[cur], compute));
            if (c + 1 < num_chunks && libB != 'X') {
                const int64_t noff = (int64_t)(c + 1) * kchunk;
                stage(1 - cur, noff, (k - noff) < kchunk ? (k - noff) : kchunk);
            }
        }
        DRV_HIP(hipStreamSynchronize(compute));
        if (libB == 'S') DRV_HIP(hipStreamSynchronize(comm));
    };

    using clock = std::chrono::steady_clock;
    auto start = clock::now();
    for (int iter = -warmup_count; iter < loop_count; ++iter) {
        if (iter == 0 && warmup_count > 0) {
            DRV_HIP(hipDeviceSynchronize());
            start = clock::now();
        }
        one_iteration();
    }
    DRV_HIP(hipDeviceSynchronize());
    const double elapsed = std::chrono::duration<double>(clock::now() - start).count();
    const double avg = elapsed / loop_count;
    std::printf("\n\nmmb myRank=0\tLibraries=%s\t RESULT: Matrix size: %d\tTime (seconds): %.9f\n\n",
                libs, matrix_size, avg);

    if (!apply_test_perturbation(d_C, mi * n)) fail("invalid MGX_TEST_PERTURB");
    int valid = 0;
    if (opt.uniform) {
        double nw = 0.0;
        valid = check_uniform(d_A, d_B, d_C, mi, n, k, k, n, n, &nw);
    } else {
        const double expected_C = 2.0 * (double)k;
        int pass = 0;
        DRV_MGX(mgx_validate_constant(d_C, mi, n, expected_C, 1.0e-9, 1.0e-12, 0, &pass, nullptr,
                                      nullptr, nullptr));
        valid = pass;
    }
    std::printf("BENCHMARK_RESULT decomposition=1D library=%s size=%d ranks=%d seconds=%.9f valid=%d "
                "loops=%d",
                libs, matrix_size, nranks, avg, valid, loop_count);
    if (extras_requested())
        std::printf(" form=%s slices=%d share_ranks=%d", form_name(opt.form), opt.slices, share);
    std::printf("\n");
    std::fflush(stdout);

    if (pipelined) {
        for (int i = 0; i < 2; ++i) {
            (void)hipEventDestroy(b_ready[i]);
            (void)hipEventDestroy(compute_done[i]);
            if (b_stage[i]) (void)hipFree(b_stage[i]);
            if (y_host[i]) mgx_host_free(y_host[i]);
        }
        (void)hipStreamDestroy(comm);
        (void)hipStreamDestroy(compute);
    }
    if (ws) (void)hipFree(ws);
    (void)hipFree(d_A);
    (void)hipFree(d_B);
    (void)hipFree(d_C);
    return valid ? 0 : 2;
}
