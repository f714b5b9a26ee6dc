// Shared host helpers of the mm_1d / mm_2d executables (one process, one GPU).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mgx.h"

namespace mgxdrv {

inline int env_int(const char *name, int fallback) {
    const char *s = std::getenv(name);
    return (s && *s) ? std::atoi(s) : fallback;
}

inline std::string env_str(const char *name, const char *fallback) {
    const char *s = std::getenv(name);
    return (s && *s) ? std::string(s) : std::string(fallback);
}

// Number of ranks the launcher started (1 when run directly).  This build
// computes one GPU's local product; the inter-rank exchange is not built.
inline int launcher_world_size() {
    const char *keys[] = {"OMPI_COMM_WORLD_SIZE", "PMI_SIZE", "PMIX_SIZE", "MV2_COMM_WORLD_SIZE",
                          "SLURM_NTASKS", "MGX_WORLD_SIZE"};
    for (const char *k : keys) {
        const char *v = std::getenv(k);
        if (v && *v) return std::atoi(v);
    }
    return 1;
}

// Device choice: "auto" -> the only device, else device 0 (one rank per
// process group of one); a number selects that device.
inline int select_device(const char *arg, int *dev_out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 3;
    int dev = std::strcmp(arg, "auto") == 0 ? 0 : std::atoi(arg);
    if (dev < 0 || dev >= count) return 3;
    if (hipSetDevice(dev) != hipSuccess) return 3;
    *dev_out = dev;
    return 0;
}

struct Options {
    int form = MGX_FORM_FP64;
    int slices = 8;
    bool uniform = false;   // MGX_INIT=uniform:SEED (else the constant A=1, B=2)
    uint64_t seed = 1;
};

// MGX_FORM=fp64|int8|bf16|mod, MGX_SLICES=<n> (slices of int8 / bf16, default 8; moduli
// of mod, default 15), MGX_INIT=const|uniform[:SEED].  Split forms also read
// MGX_SPLIT_K (columns per split product in the panel / chunk loops, whole
// panels or chunks; default MGX_SPLIT_GROUP_K) where the loop runs.
inline bool parse_options(Options *o) {
    const std::string form = env_str("MGX_FORM", "fp64");
    if (form == "fp64") o->form = MGX_FORM_FP64;
    else if (form == "int8") o->form = MGX_FORM_INT8;
    else if (form == "mod") o->form = MGX_FORM_INT8_MOD;
    else if (form == "bf16") o->form = MGX_FORM_BF16;
    else return false;
    o->slices = env_int("MGX_SLICES", o->form == MGX_FORM_INT8_MOD ? 15 : 8);
    const std::string init = env_str("MGX_INIT", "const");
    if (init == "const") {
        o->uniform = false;
    } else if (init.rfind("uniform", 0) == 0) {
        o->uniform = true;
        const size_t colon = init.find(':');
        if (colon != std::string::npos) o->seed = std::strtoull(init.c_str() + colon + 1, nullptr, 10);
    } else {
        return false;
    }
    return true;
}

// True when any knob of this build that changes the workload or its timing is
// set.  Only then do the executables append their extra keys (form, slices,
// share, panels) to the BENCHMARK_RESULT line; with none set the line is the
// reference's format string byte for byte (1D-row-block/mmb.cu:964,
// 2D-block-matrix/mmb.cu:398).
inline bool extras_requested() {
    const char *knobs[] = {"MGX_FORM", "MGX_SLICES", "MGX_INIT", "MGX_SHARE_RANKS",
                           "MGX_SHARE_GRID", "MGX_SPLIT_K", "MGX_SUMMA_GRAPH",
                           "MGX_HOST_PINNED", "MGX_SUMMA_SERIALIZE", "MGX_RESULT_EXTRAS"};
    for (const char *k : knobs) {
        const char *v = std::getenv(k);
        if (v && *v) return true;
    }
    return false;
}

// Test hook MGX_TEST_PERTURB=<index>:<value>: overwrite element `index` of the
// `count` contiguous doubles of the device result C with `value` (strtod
// syntax, so "nan" and "inf" work) after the timed region and before
// validation, so the failing-validation path can be exercised end to end.
// Returns false on a malformed value or an index outside C.
inline bool apply_test_perturbation(double *dC, int64_t count) {
    const char *s = std::getenv("MGX_TEST_PERTURB");
    if (!s || !*s) return true;
    char *end = nullptr;
    const long long idx = std::strtoll(s, &end, 10);
    if (end == s || *end != ':' || idx < 0 || idx >= count) return false;
    const char *vs = end + 1;
    const double value = std::strtod(vs, &end);
    if (end == vs || *end != '\0') return false;
    return hipMemcpy(dC + idx, &value, sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
}

inline const char *form_name(int form) {
    return form == MGX_FORM_INT8       ? "int8"
           : form == MGX_FORM_INT8_MOD ? "mod"
           : form == MGX_FORM_BF16     ? "bf16"
                                       : "fp64";
}

// splitmix64 counter stream, the definition used by mgx_fill.
inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uniform_at(uint64_t seed, uint64_t idx) {
    return 2.0 * (double)(splitmix64_at(seed, idx + 1) >> 11) * 0x1.0p-53 - 1.0;
}

// Error of a uniform-input run against the GPU double-double check on an even
// sample of rows; valid when the normwise error is within the FP64 gate.
inline int check_uniform(const double *dA, const double *dB, const double *dC, int64_t m,
                         int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc,
                         double *normwise) {
    const int64_t want = m < 16 ? m : 16;
    int64_t rows[16];
    for (int64_t i = 0; i < want; ++i) rows[i] = (i * m) / want;
    mgx_error_stats st;
    if (mgx_error_report(dA, dB, dC, m, n, k, lda, ldb, ldc, rows, want, &st) != MGX_OK) {
        std::fprintf(stderr, "accuracy check failed: %s\n", mgx_last_error());
        return 0;
    }
    *normwise = st.normwise;
    std::printf("Accuracy (sampled %lld rows): normwise %.3e componentwise %.3e\n",
                (long long)st.rows_checked, st.normwise, st.componentwise);
    return st.normwise <= 1e-13 ? 1 : 0;
}

}  // namespace mgxdrv
