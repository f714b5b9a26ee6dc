// Internal helpers shared by the libmgx HIP sources (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "../../include/mgx.h"

namespace mgx {

// Records `msg` as the calling thread's last error and returns `status`.
int set_error(int status, const std::string &msg);
void clear_error();

#define MGX_HIP_TRY(expr)                                                          \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::mgx::set_error(MGX_ERR_HIP, std::string(#expr " failed: ") +  \
                                                   hipGetErrorString(_e));         \
    } while (0)

#define MGX_ARG_CHECK(cond, msg)                                                   \
    do {                                                                           \
        if (!(cond)) return ::mgx::set_error(MGX_ERR_INVALID_ARG, (msg));          \
    } while (0)

// Checks the launch status of the kernel just enqueued.
inline int launch_status(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(MGX_ERR_HIP, std::string(what) + " launch failed: " + hipGetErrorString(e));
    return MGX_OK;
}

// Per-device library-owned device buffer, grown on demand and kept until exit.
// `stream` is the stream the caller's work on the buffer will be queued on:
// when the previous user was another stream, that stream's last use is
// waited for first; growing it inside a stream capture, or switching streams
// inside one, fails with MGX_ERR_WORKSPACE (the caller must pass a workspace).
// Every successful acquire is followed by scratch_release(slot, stream) once
// the work has been queued.
void *device_scratch(int slot, size_t bytes, void *stream, int *status);
int scratch_release(int slot, void *stream);

enum ScratchSlot { SCRATCH_VALIDATE = 0, SCRATCH_SPLIT = 1, SCRATCH_MAX = 2 };

// splitmix64: state advances by the golden gamma; output i (1-based) of a
// generator seeded with `seed` is mix(seed + i * gamma).
__host__ __device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform[-1,1): 2 * (x >> 11) * 2^-53 - 1 (exact in FP64).
__host__ __device__ inline double u64_to_uniform(uint64_t x) {
    return 2.0 * (double)(x >> 11) * 0x1.0p-53 - 1.0;
}

}  // namespace mgx
