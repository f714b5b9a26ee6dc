/*
 * mgx.h -- C ABI of libmgx, the MI355X (gfx950) local FP64 block product.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status (MGX_OK == 0) instead of aborting the process.
 * Matrices are row-major with explicit leading dimensions (in elements).
 * Device pointers are owned by the caller unless stated otherwise.
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *
 * Threading: entry points may be called from any host thread; calls that
 * share a stream are ordered by that stream.  The library keeps per-device
 * scratch allocations (validation partials, split-form workspaces grown on
 * demand) that are reused across calls on the same device.  A call that uses
 * the library-owned split workspace on another stream than the previous such
 * call waits (hipStreamWaitEvent) for that call's work first, so calls issued
 * from one host thread are ordered as issued; calls from several host threads
 * at once must bring their own workspaces.  Inside a stream capture the
 * library scratch can neither grow nor change streams: such a call returns
 * MGX_ERR_WORKSPACE (pass a workspace instead).
 */
#ifndef MGX_H
#define MGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    MGX_OK = 0,
    MGX_ERR_INVALID_ARG = 1,   /* a size, leading dimension or pointer is invalid */
    MGX_ERR_HIP = 2,           /* a HIP runtime call failed (see mgx_last_error) */
    MGX_ERR_UNSUPPORTED = 3,   /* form / slice count / option not supported */
    MGX_ERR_WORKSPACE = 4,     /* caller workspace too small, or library scratch not usable
                                  inside a stream capture */
    MGX_ERR_ALLOC = 5          /* device or pinned allocation failed */
};

/* Split-precision forms. */
enum {
    MGX_FORM_FP64 = 0,         /* plain FP64 MFMA (v_mfma_f64_16x16x4_f64) */
    MGX_FORM_INT8 = 1,         /* Ozaki scheme I: INT8 digit slices, exact INT32 MFMA sums of
                                  the s(s+1)/2 slice pairs, FP64 folds */
    MGX_FORM_BF16 = 2,         /* the INT8 form's digit slices held as bf16, products on the
                                  bf16 MFMA with FP32 sums moved to INT32 while exact: same
                                  result bit for bit as MGX_FORM_INT8, half its matrix rate */
    MGX_FORM_INT8_MOD = 3      /* Ozaki scheme II: INT8 residues of the 2^beta-scaled integer
                                  operands modulo `slices` coprime moduli, one exact INT32 MFMA
                                  product per modulus, Chinese remaindering to FP64 */
};

/* Block-tile variants of the product kernels (autotuner knob). */
enum {
    MGX_TILE_DEFAULT = 0,   /* the large tile */
    MGX_TILE_LARGE = 1,     /* FP64: 128x128 C block per workgroup; INT8: 256x256 */
    MGX_TILE_SMALL = 2      /* FP64: 64x64; INT8: 128x128 */
};

/* Human-readable text of the last error on the calling thread. */
const char *mgx_last_error(void);
/* Library version string, e.g. "0.1.0". */
const char *mgx_version(void);

/* Published dense peak of the MI355X resource named by `form`, in TFLOP/s
 * (TOP/s for INT8): FP64 MFMA 78.6, INT8 5000, BF16 2500; form = -1 gives
 * HBM bandwidth in GB/s (8000). Returns a negative value for unknown forms. */
double mgx_peak(int form);

/* ---- plain FP64 GEMM ---------------------------------------------------
 * Replaces ABMultiply (1D-row-block/mulmat_kernel.cu:86-92,
 * 2D-block-matrix/mulmat_kernel.cu:81-87):  C[m x n] = A[m x k] . B[k x n].
 * `w` is the reference's tile width argument; it must be positive and is
 * otherwise ignored (the tile is chosen by the library).
 * Asynchronous on `stream`. */
int mgx_dgemm(const double *A, const double *B, double *C,
              int64_t m, int64_t n, int64_t k,
              int64_t lda, int64_t ldb, int64_t ldc,
              int w, void *stream);

/* Replaces ABMultiplyAccumulateAsync (1D-row-block/mulmat_kernel.cu:148-158,
 * 2D-block-matrix/mulmat_kernel.cu:142-151):
 *   C += A[:, k_offset : k_offset + k_chunk] . B_chunk
 * B_chunk is a packed k_chunk x n block (ldb).  Columns of A with
 * k_offset + kk >= k_total contribute nothing (the reference loads them as
 * zero).  The chunk partial product is formed first and then added to C
 * once, as in the reference.  Asynchronous on `stream`. */
int mgx_dgemm_accumulate(const double *A, const double *B_chunk, double *C,
                         int64_t m, int64_t n, int64_t k_total,
                         int64_t k_offset, int64_t k_chunk,
                         int64_t lda, int64_t ldb, int64_t ldc,
                         int w, void *stream);

/* C = A.B (accumulate == 0) or C += A.B with an explicit tile variant.
 * No counterpart in the reference (its tile is fixed at 32); used by the
 * autotuner. */
int mgx_dgemm_ex(int tile, int accumulate, const double *A, const double *B, double *C,
                 int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                 void *stream);

/* The panel-fused FP64 product: C = (accumulate == 0) or C += the sum over
 * panels p of A[:, p] . B[p, :], panels of width `panel` along k (ragged last
 * panel), each panel's partial formed from zero and added into C in panel
 * order -- the result of one ABMultiplyAccumulateAsync call per panel
 * (2D-block-matrix/mmb.cu:356-358), bit for bit, in one launch that reads and
 * writes C once.  No counterpart in the reference: it fuses the reference's
 * per-panel launch sequence; mgx_dgemm is the case panel >= k.  `tile` is a
 * MGX_TILE_* variant.  Asynchronous on `stream`. */
int mgx_dgemm_panels(int tile, int accumulate, const double *A, const double *B, double *C,
                     int64_t m, int64_t n, int64_t k, int64_t panel,
                     int64_t lda, int64_t ldb, int64_t ldc, void *stream);

/* ---- constant-result validation ----------------------------------------
 * Replaces validate_matrix_C (1D-row-block/mulmat_kernel.cu:228-326).
 * An element is invalid iff |C-expected| > abs_tol AND |C-expected|/|expected|
 * > rel_tol (the relative error is 0 when expected == 0), exactly the
 * reference rule (1D-row-block/mulmat_kernel.cu:180-187).  Synchronous.
 * Prints the "NUMERICAL VALIDATION" block to stdout when rank == 0.
 * Returns MGX_OK on a completed check; *pass_out = 1 if no element is
 * invalid.  Any of the out pointers may be NULL. */
int mgx_validate_constant(const double *dC, int64_t rows, int64_t cols,
                          double expected, double abs_tol, double rel_tol,
                          int rank, int *pass_out, uint64_t *invalid_out,
                          double *max_abs_out, double *max_rel_out);

/* ---- split-precision (Ozaki) GEMM ---------------------------------------
 * C = A.B (accumulate == 0) or C += A.B (accumulate != 0), FP64 in and out,
 * computed from `slices` INT8 (or BF16) slices of the power-of-two scaled
 * rows of A and columns of B.  Slice-pair products run on the INT8 (BF16)
 * MFMA cores with exact INT32 (FP32) accumulation; FP64 is used for the
 * exponent scan, the slice extraction and the reconstruction into C.
 * No counterpart in the reference: it is a second implementation of the
 * same local product (ABMultiply / ABMultiplyAccumulateAsync semantics).
 *
 * Workspace: `workspace` (device, 256-byte aligned) of `workspace_bytes`
 * >= mgx_split_workspace_bytes(...); pass workspace == NULL to use a
 * library-owned per-device buffer grown on demand.
 * Inputs must be finite; a row of A or column of B holding Inf/NaN
 * produces NaN in the corresponding row/column of C.
 * Size limits (grid dimensions of the operand passes): m and k at most
 * 4194048; larger calls return MGX_ERR_INVALID_ARG before any launch.
 * MGX_FORM_INT8_MOD: `slices` is the number of moduli (2..16); the result
 * is ldexp(RN(A'B'), e_i - beta_i + f_j - delta_j) with A', B' the integer
 * operands, row i of A scaled by 2^(beta_i - e_i) and column j of B by
 * 2^(delta_j - f_j) (per-line exponents, see mgx_split_mod_beta), so each
 * output is rounded once.
 * Its events: [2] before the residue products, [3] after them (the
 * reconstruction into C follows [3]).
 * MGX_FORM_BF16: `slices` 1..9 as for MGX_FORM_INT8, same digits, group
 * sums, chunks and folds, so the same result bit for bit; always runs its
 * own 256 x 128 block tile (it keeps FP32 and INT32 accumulators) and
 * ignores `tile`. */
size_t mgx_split_workspace_bytes(int form, int slices,
                                 int64_t m, int64_t n, int64_t k);
int mgx_split_gemm(int form, int slices,
                   const double *A, const double *B, double *C,
                   int64_t m, int64_t n, int64_t k,
                   int64_t lda, int64_t ldb, int64_t ldc,
                   int accumulate, void *workspace, size_t workspace_bytes,
                   void *stream);

/* mgx_split_gemm that also records four caller-created hipEvent_t handles
 * (passed as void*; NULL entries are skipped) on `stream` at the phase
 * boundaries: [0] before the exponent scan, [1] before slice extraction,
 * [2] before the slice-pair product kernel, [3] after it.  Asynchronous;
 * the caller reads hipEventElapsedTime afterwards.  Used to time the
 * product kernel inside a timed region without synchronising. */
int mgx_split_gemm_events(int form, int slices,
                          const double *A, const double *B, double *C,
                          int64_t m, int64_t n, int64_t k,
                          int64_t lda, int64_t ldb, int64_t ldc,
                          int accumulate, void *workspace, size_t workspace_bytes,
                          void *stream, void *const *events);

/* mgx_split_gemm with an explicit tile variant (MGX_TILE_*). */
int mgx_split_gemm_ex(int form, int slices, int tile,
                      const double *A, const double *B, double *C,
                      int64_t m, int64_t n, int64_t k,
                      int64_t lda, int64_t ldb, int64_t ldc,
                      int accumulate, void *workspace, size_t workspace_bytes,
                      void *stream);

/* No counterpart in the reference: mgx_split_gemm for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid allocation).  It calls no hipMalloc / hipFree and
 * takes no workspace: the workspace is a library-owned static device array of
 * MGX_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Calls
 * whose mgx_split_workspace_bytes(form, slices, m, n, k) exceeds it return
 * MGX_ERR_WORKSPACE before any launch (for example, 15 moduli fit up to
 * 1024 x 1024 x 1024: 45 MiB).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other host threads) race.
 * Otherwise as mgx_split_gemm, with the same result bit for bit. */
#define MGX_NOALLOC_SCRATCH_BYTES (64u << 20)
int mgx_split_gemm_noalloc(int form, int slices,
                           const double *A, const double *B, double *C,
                           int64_t m, int64_t n, int64_t k,
                           int64_t lda, int64_t ldb, int64_t ldc,
                           int accumulate, void *stream);

/* Supported slice-count range for a split form (inclusive); for
 * MGX_FORM_INT8_MOD the count is the number of moduli. */
int mgx_split_slices_range(int form, int *min_slices, int *max_slices);

/* Guaranteed scaling exponent of MGX_FORM_INT8_MOD for `moduli` moduli and
 * depth k.  Each row of A and column of B is quantised to an integer line of
 * at most 2^beta (the form's only error), with beta chosen per line from an
 * exponent histogram of the line so that a Cauchy-Schwarz bound keeps every
 * |A'B'| entry below m_0 ... m_{n-1} / 2; lines spread over many binades get
 * more bits.  This returns the exponent a line gets when all its elements
 * sit in its top binade, which every line gets at least (the classical
 * largest beta <= 53 with 2 k 4^beta < m_0 ... m_{n-1}, up to a rounding
 * margin).  Returns -1 when the moduli cannot hold the product (the call
 * would fail). */
int mgx_split_mod_beta(int moduli, int64_t k);

/* ---- accuracy report ----------------------------------------------------
 * Error of a computed device product C (m x n) against a double-double
 * reference of A.B evaluated on the GPU for the listed rows (host array of
 * `nrows` row indices): normwise = ||C - AB||_F / ||AB||_F over those rows,
 * componentwise = max |C - AB| / (|A||B|).  A NaN in C reports Inf.
 * Synchronous.  No counterpart in the reference (its only check is the
 * constant known answer); it backs the error statement of every result. */
typedef struct mgx_error_stats {
    double normwise;
    double componentwise;
    double max_abs;
    int64_t rows_checked;
} mgx_error_stats;

int mgx_error_report(const double *A, const double *B, const double *C,
                     int64_t m, int64_t n, int64_t k,
                     int64_t lda, int64_t ldb, int64_t ldc,
                     const int64_t *rows, int64_t nrows, mgx_error_stats *out);

/* ---- SUMMA panel streaming ----------------------------------------------
 * Replaces the local part of the 2D SUMMA panel loop
 * (2D-block-matrix/mmb.cu:239-371): for panels of width `panel` (ragged last
 * panel), copy the m x kc column panel of the host A strip and the kc x n
 * row panel of the host B strip to device double buffers on a transfer
 * stream and run C += A_p . B_p on a compute stream, with event handoff
 * (panel ready / compute done) in both directions.  The per-panel
 * MPI/NCCL/NVSHMEM broadcasts of the reference are replaced by these
 * host-to-device copies (the single-GPU stand-in).
 * Host strips: pinned (mgx_host_alloc) strips are gathered directly by
 * strided DMA, one A copy and one B copy per group; pageable strips are
 * packed (the reference's pack_a / pack_b) by a pool of host_threads host
 * threads into two library-owned pinned staging buffers in units of 512
 * columns (parts of wider panels), packing the next unit while the current
 * one is copied into its place in the group's device buffer; the staging
 * buffers are written once before the timed loop.  dC is a device m x n matrix that is accumulated into (zero it for
 * C = A.B).  Synchronous: all resources (buffers, streams, host threads) are
 * created before and released after the timed loop.
 * Consecutive panels form groups (mgx_summa_groups) and each group is one
 * launch: FP64 by mgx_dgemm_panels (one partial per panel, the reference's
 * result), split forms by one split product C += A_g . B_g, so the exponent
 * scan, operand conversion and reconstruction are paid per group, not per
 * panel.  The copy of a group overlaps the product of the previous one. */
enum {
    MGX_SUMMA_COPY_ASYNC = 0,      /* copies overlap compute (MMM, YYY, NNN codes) */
    MGX_SUMMA_COPY_HOST_SYNC = 1   /* host waits for each panel copy (CCC, XXX, WWW, SSS) */
};

typedef struct mgx_summa_params {
    int64_t m, n, k;    /* local C is m x n; reduction length k */
    int64_t panel;      /* panel width (the reference's MGXBENCH_SUMMA_BK) */
    int form;           /* MGX_FORM_FP64, MGX_FORM_INT8 or MGX_FORM_INT8_MOD */
    int slices;         /* slices (INT8) or moduli (INT8_MOD) */
    int copy_mode;      /* MGX_SUMMA_COPY_* */
    int use_graph;      /* capture the whole panel loop in one hipGraph (pinned strips) */
    int serialize;      /* debug: synchronise after every copy and every product */
    int64_t split_k;    /* group schedule (mgx_summa_groups): > 0 uniform groups of
                           this many columns (whole panels), 0 the default width,
                           < 0 the tapered schedule on groups of -split_k */
    int host_threads;   /* threads packing pageable strips, 1..64; 0 = automatic
                           (min(16, hardware threads)); unused with pinned strips */
    int tile;           /* MGX_TILE_* block tile of the group products */
} mgx_summa_params;

typedef struct mgx_summa_stats {
    double seconds;     /* host wall time of the panel loop (launch to completion) */
    int64_t panels;     /* number of panels, including a ragged last one */
    int host_pinned;    /* 1 if the host strips were pinned (no host packing) */
    int host_threads;   /* host threads that packed pageable strips (1 when pinned) */
    int64_t groups;     /* launches (panel groups) of the loop */
} mgx_summa_stats;

/* Default reduction columns per launch in a panel loop, split products and
 * FP64 panel groups.  A streamed loop ends with its last group's product, and
 * narrow groups keep that short; their fixed costs (split forms: exponent
 * scan and reconstruction) hide under the next group's copy.  Measured on
 * the 8192 x 8192 x 16384 share with pinned strips: 1024 columns take
 * 40.1 ms (15 moduli) and 40.2 ms (FP64), against 43.1 and 42.0 ms with the
 * 4096 / 2048 groups. */
#define MGX_SPLIT_GROUP_K 1024
#define MGX_FP64_GROUP_K 1024

/* Group width of uniform panel groups: the smallest whole number of panels
 * covering split_k columns (MGX_SPLIT_GROUP_K, or MGX_FP64_GROUP_K for
 * MGX_FORM_FP64, when split_k <= 0), capped at k.  Returns -1 for
 * non-positive panel or k, or an unknown form. */
int64_t mgx_split_group_k(int form, int64_t panel, int64_t split_k, int64_t k);

/* The group schedule of mgx_summa_run: column widths of its launches, in
 * order (sum k).  split_k >= 0: uniform groups of
 * mgx_split_group_k(form, panel, split_k, k) columns, the last one shorter.
 * split_k < 0, the tapered schedule: groups of G = mgx_split_group_k(form,
 * panel, -split_k, k) columns (P panels), except that when k > G the loop
 * ends with T = min(4, P / t) groups of t = max(1, P / 4) panels, so the
 * product left after the last copy is short (worth it for wide groups).
 * The ragged last panel ends the last group.  Writes min(count, capacity) widths (widths may be NULL) and
 * returns the group count, or -1 as mgx_split_group_k. */
int64_t mgx_summa_groups(int form, int64_t panel, int64_t split_k, int64_t k,
                         int64_t *widths, int64_t capacity);

int mgx_summa_run(const mgx_summa_params *params,
                  const double *hA, int64_t lda,
                  const double *hB, int64_t ldb,
                  double *dC, int64_t ldc,
                  mgx_summa_stats *stats);

/* Pinned host memory for SUMMA strips (hipHostMalloc / hipHostFree). */
int mgx_host_alloc(void **ptr, size_t bytes);
int mgx_host_free(void *ptr);

/* ---- autotuner measurement ------------------------------------------------
 * One candidate configuration of the local product: form (MGX_FORM_FP64 or
 * MGX_FORM_INT8), slices (INT8), tile (MGX_TILE_*) and SUMMA panel width.
 * mgx_time_local_product runs the compute side of the SUMMA panel loop on
 * device-resident operands: C = sum over groups g of A[:, g] . B[g, :], with
 * groups of mgx_split_group_k(form, panel, split_k, k) columns, each one
 * launch (FP64: mgx_dgemm_panels, one partial per panel); the first group
 * overwrites C, later groups accumulate.  `warmup` untimed then `reps` timed repetitions; returns the
 * mean milliseconds per full product (hipEvent timing).  workspace may be
 * NULL (library scratch). */
typedef struct mgx_tune_config {
    int form;
    int slices;
    int tile;
    int64_t panel;
    int64_t split_k;    /* as mgx_summa_params.split_k */
} mgx_tune_config;

int mgx_time_local_product(const mgx_tune_config *cfg,
                           const double *A, const double *B, double *C,
                           int64_t m, int64_t n, int64_t k,
                           int64_t lda, int64_t ldb, int64_t ldc,
                           void *workspace, size_t workspace_bytes,
                           int warmup, int reps, double *ms_out);

/* ---- matrix-core ceiling probe ---------------------------------------------
 * No counterpart in the reference: a measurement aid for the product
 * kernel's roofline.  Launches `blocks` workgroups of 8 waves that issue
 * only independent v_mfma_i32_16x16x64_i8 (8 x 4 accumulator fragments per
 * wave, as in the product kernel), `iters` (even) rounds of 32 per wave.
 * random_operands != 0: every round re-reads the wave's 12 fragments from a
 * 64 KiB LDS buffer of pseudo-random bytes; 0: every operand byte is 1
 * (register operands) and every out[t] equals 8192 * iters.  out holds
 * blocks * 512 int32 (one per thread).
 * Asynchronous on `stream`.  mgx_mfma_i8_ceiling_ops gives the INT8
 * operations of one launch. */
int mgx_mfma_i8_ceiling(int32_t *out, int blocks, int iters, int random_operands,
                        void *stream);
double mgx_mfma_i8_ceiling_ops(int blocks, int iters);

/* ---- learned cost model ---------------------------------------------------
 * A 3-layer, width-64 FP32 MLP (in_dim -> 64 -> 64 -> 1, ReLU) trained and
 * evaluated on the GPU by the library's own kernels.  It replaces the fitted
 * runtime models of the reference autotuner (autotuning/predictive.py:100-233:
 * power law, lstsq "physical" fit, ridge "ml" fit).  Parameters are one FP32
 * vector: W1 [in_dim][64], b1 [64], W2 [64][64], b2 [64], W3 [64], b3.
 * All arrays are host arrays; calls are synchronous. */
#define MLP_MAX_IN 24
#define MLP_HIDDEN 64
int64_t mgx_mlp_param_count(int in_dim);                 /* -1 if in_dim is out of range */
int mgx_mlp_init(float *theta, int in_dim, uint64_t seed);  /* He-uniform weights, zero biases */
int mgx_mlp_predict(const float *theta, int in_dim, const float *X, int64_t n, float *y);
/* Mean squared error over the n samples and its gradient. */
int mgx_mlp_gradient(const float *theta, int in_dim, const float *X, const float *t,
                     int64_t n, float *grad, float *loss);
/* `epochs` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8, decoupled-free
 * L2 `weight_decay` added to the gradient); theta is updated in place. */
int mgx_mlp_train(float *theta, int in_dim, const float *X, const float *t, int64_t n,
                  int epochs, float lr, float weight_decay, float *loss_history);

/* ---- synthetic inputs ---------------------------------------------------
 * Fill a rows x cols row-major device matrix (ld) on `stream`.
 * kind 0: constant `value` (the reference init: A = 1, B = 2,
 *         1D-row-block/mmb.cu:571-605, 2D-block-matrix/mmb.cu:151-154).
 * kind 1: uniform[-1,1): element (r,c) = 2*(x>>11)*2^-53 - 1 with
 *         x = the (r*cols+c+1)-th output of splitmix64 seeded with `seed`.
 * kind 2: accuracy stress: element i = r*cols+c takes x = output 2i+1 and
 *         y = output 2i+2 of the same generator; value =
 *         (y&1 ? -1 : 1) * (1 + (x>>12)*2^-52) * 2^((y>>11) % 61 - 30).
 *         The row/column scaling of the stress inputs is applied separately
 *         (mgx_fill_exponents + mgx_scale_rows_cols). */
int mgx_fill(double *X, int64_t rows, int64_t cols, int64_t ld,
             int kind, uint64_t seed, double value, void *stream);

/* e[i] = lo + ((output i+1 of splitmix64(seed)) >> 11) % (hi-lo+1);
 * e is a device int32 array. */
int mgx_fill_exponents(int32_t *e, int64_t count, uint64_t seed, int lo, int hi,
                       void *stream);

/* X[r][c] *= 2^(row_exp[r] + col_exp[c]); either exponent array may be NULL.
 * Exponent arrays are device int32 arrays. */
int mgx_scale_rows_cols(double *X, int64_t rows, int64_t cols, int64_t ld,
                        const int32_t *row_exp, const int32_t *col_exp,
                        void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MGX_H */
