// Matrix-core ceiling probe: the INT8 MFMA rate this GPU sustains when the
// instruction stream is nothing but independent v_mfma_i32_16x16x64_i8 fed
// from LDS (or registers), laid out like the product kernel (8 waves per
// workgroup, two per SIMD, 8 x 4 accumulator fragments per wave).  With
// random operands the rate is set by the part's power management, not by
// the published dense peak; the product kernel's roofline fraction is read
// against both.
#include "mgx_internal.hpp"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int CFI = 8, CFJ = 4;

// RANDOM: every round re-reads all 12 operand fragments of each wave from a
// 64 KiB LDS buffer of pseudo-random bytes (four 12 KiB regions in turn),
// as an LDS-fed GEMM main loop does, minus global loads and barriers.
// Otherwise every operand byte is 1 (register operands), so each
// accumulator element grows by 64 per MFMA and out[t] = 32 * 4 * 64 * iters.
template <bool RANDOM>
__global__ __launch_bounds__(512, 1) void mfma_i8_ceiling_kernel(int32_t *out, int iters) {
    __shared__ __attribute__((aligned(16))) int8_t lds[RANDOM ? 65536 : 16];
    const uint32_t t = threadIdx.x + blockIdx.x * 512u;
    v4i acc[CFI][CFJ];
#pragma unroll
    for (int i = 0; i < CFI; ++i)
#pragma unroll
        for (int j = 0; j < CFJ; ++j) acc[i][j] = v4i{0, 0, 0, 0};
    if constexpr (RANDOM) {
        uint32_t h = t * 2654435761u + 12345u;
        for (int i = threadIdx.x; i < 65536 / 4; i += 512) {
            h ^= h << 13;
            h ^= h >> 17;
            h ^= h << 5;
            reinterpret_cast<uint32_t *>(lds)[i] = h;
        }
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int it = 0; it < iters; ++it) {
            // region (it + wave) % 4; 16 bytes per lane per fragment
            const int base = ((it + wave) & 3) * 12288 + lane * 16;
            v4i fa[CFI], fb[CFJ];
#pragma unroll
            for (int i = 0; i < CFI; ++i) fa[i] = *(const v4i *)(lds + base + i * 1024);
#pragma unroll
            for (int j = 0; j < CFJ; ++j) fb[j] = *(const v4i *)(lds + base + (CFI + j) * 1024);
#pragma unroll
            for (int i = 0; i < CFI; ++i)
#pragma unroll
                for (int j = 0; j < CFJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0,
                                                                       0, 0);
        }
    } else {
        const v4i one = {0x01010101, 0x01010101, 0x01010101, 0x01010101};
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int i = 0; i < CFI; ++i)
#pragma unroll
                for (int j = 0; j < CFJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(one, one, acc[i][j], 0, 0, 0);
        }
    }
    int32_t s = 0;
#pragma unroll
    for (int i = 0; i < CFI; ++i)
#pragma unroll
        for (int j = 0; j < CFJ; ++j) s += acc[i][j][0] + acc[i][j][1] + acc[i][j][2] + acc[i][j][3];
    out[t] = s;
}

}  // namespace

extern "C" int mgx_mfma_i8_ceiling(int32_t *out, int blocks, int iters, int random_operands,
                                   void *stream) {
    mgx::clear_error();
    MGX_ARG_CHECK(out != nullptr, "out is NULL");
    MGX_ARG_CHECK(blocks > 0 && iters > 0 && iters % 2 == 0, "blocks > 0, iters > 0 and even");
    hipStream_t st = (hipStream_t)stream;
    if (random_operands)
        hipLaunchKernelGGL(mfma_i8_ceiling_kernel<true>, dim3((unsigned)blocks), dim3(512), 0, st,
                           out, iters);
    else
        hipLaunchKernelGGL(mfma_i8_ceiling_kernel<false>, dim3((unsigned)blocks), dim3(512), 0, st,
                           out, iters);
    return mgx::launch_status("mfma_i8_ceiling_kernel");
}

extern "C" double mgx_mfma_i8_ceiling_ops(int blocks, int iters) {
    // 8 waves x 32 MFMAs per iteration x 16 x 16 x 64 multiply-adds x 2
    return (double)blocks * 8.0 * (double)iters * CFI * CFJ * 16.0 * 16.0 * 64.0 * 2.0;
}
