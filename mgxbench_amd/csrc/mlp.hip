// Learned cost model: a 3-layer, width-64 FP32 MLP trained and evaluated on
// the GPU by hand-written kernels.
//
//   h1 = relu(x W1 + b1)      x: D features (D <= MLP_MAX_IN)
//   h2 = relu(h1 W2 + b2)     hidden width H = 64
//   y  = h2 W3 + b3           scalar prediction (the autotuner uses log time)
//   loss = mean over samples of (y - t)^2
//
// Parameters live in one FP32 vector, in this order: W1 [D][H] row-major,
// b1 [H], W2 [H][H] row-major, b2 [H], W3 [H], b3 [1].
//
// One wave64 per workgroup; lane j owns hidden unit j of both layers.  A
// workgroup walks a contiguous chunk of samples, keeps its gradient
// contributions in registers and writes one partial gradient vector; a second
// kernel sums the partials in chunk order, so gradients are deterministic.
// Adam (Kingma & Ba) updates the parameters in place.
#include "mgx_internal.hpp"

#include <cmath>
#include <vector>

namespace {

constexpr int H = 64;
constexpr int MAX_IN = MLP_MAX_IN;

__host__ __device__ inline int64_t n_params(int D) { return (int64_t)D * H + H + H * H + H + H + 1; }

struct Offsets {
    int w1, b1, w2, b2, w3, b3;
};
__host__ __device__ inline Offsets offsets(int D) {
    Offsets o;
    o.w1 = 0;
    o.b1 = D * H;
    o.w2 = o.b1 + H;
    o.b2 = o.w2 + H * H;
    o.w3 = o.b2 + H;
    o.b3 = o.w3 + H;
    return o;
}

__device__ inline float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Forward pass for samples [0, n): y[s].
__global__ __launch_bounds__(64) void mlp_forward_kernel(const float *__restrict__ theta,
                                                         const float *__restrict__ X, int64_t n,
                                                         int D, float *__restrict__ y) {
    __shared__ float w2s[H * H];
    __shared__ float h1s[H];
    __shared__ float xs[MAX_IN];
    const Offsets o = offsets(D);
    const int j = threadIdx.x;
    for (int i = j; i < H * H; i += 64) w2s[i] = theta[o.w2 + i];
    __syncthreads();
    for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
        if (j < D) xs[j] = X[s * D + j];
        __syncthreads();
        float a1 = theta[o.b1 + j];
        for (int d = 0; d < D; ++d) a1 = fmaf(xs[d], theta[o.w1 + d * H + j], a1);
        h1s[j] = fmaxf(a1, 0.0f);
        __syncthreads();
        float a2 = theta[o.b2 + j];
        for (int i = 0; i < H; ++i) a2 = fmaf(h1s[i], w2s[i * H + j], a2);
        const float h2 = fmaxf(a2, 0.0f);
        const float out = wave_sum(h2 * theta[o.w3 + j]) + theta[o.b3];
        if (j == 0) y[s] = out;
        __syncthreads();
    }
}

// Gradient partials of the mean squared error over samples [c*chunk, ...)
// of chunk c; `scale` = 1 / (total sample count).
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float *__restrict__ theta,
                                                      const float *__restrict__ X,
                                                      const float *__restrict__ t, int64_t n,
                                                      int D, int64_t chunk, float scale,
                                                      float *__restrict__ partials,
                                                      float *__restrict__ loss_partials) {
    __shared__ float w2s[H * H];
    __shared__ float h1s[H];
    __shared__ float dh2s[H];
    __shared__ float xs[MAX_IN];
    const Offsets o = offsets(D);
    const int64_t P = n_params(D);
    const int j = threadIdx.x;
    for (int i = j; i < H * H; i += 64) w2s[i] = theta[o.w2 + i];
    __syncthreads();

    float gw1[MAX_IN];
    float gw2[H];
#pragma unroll
    for (int d = 0; d < MAX_IN; ++d) gw1[d] = 0.0f;
#pragma unroll
    for (int i = 0; i < H; ++i) gw2[i] = 0.0f;
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
    const float w3j = theta[o.w3 + j];

    const int64_t s0 = blockIdx.x * chunk;
    const int64_t s1 = min(n, s0 + chunk);
    for (int64_t s = s0; s < s1; ++s) {
        if (j < D) xs[j] = X[s * D + j];
        __syncthreads();
        float a1 = theta[o.b1 + j];
        for (int d = 0; d < D; ++d) a1 = fmaf(xs[d], theta[o.w1 + d * H + j], a1);
        const float h1 = fmaxf(a1, 0.0f);
        h1s[j] = h1;
        __syncthreads();
        float a2 = theta[o.b2 + j];
        for (int i = 0; i < H; ++i) a2 = fmaf(h1s[i], w2s[i * H + j], a2);
        const float h2 = fmaxf(a2, 0.0f);
        const float y = wave_sum(h2 * w3j) + theta[o.b3];
        const float diff = y - t[s];
        loss = fmaf(diff * diff, scale, loss);
        const float dy = 2.0f * diff * scale;
        gw3 = fmaf(h2, dy, gw3);
        gb3 += dy;
        const float dh2 = a2 > 0.0f ? w3j * dy : 0.0f;
        dh2s[j] = dh2;
        gb2 += dh2;
        // gW2[i][j] += h1_i * dh2_j (lane j owns column j)
#pragma unroll
        for (int i = 0; i < H; ++i) gw2[i] = fmaf(h1s[i], dh2, gw2[i]);
        __syncthreads();
        float back = 0.0f;
        for (int k = 0; k < H; ++k) back = fmaf(w2s[j * H + k], dh2s[k], back);
        const float dh1 = a1 > 0.0f ? back : 0.0f;
        gb1 += dh1;
#pragma unroll
        for (int d = 0; d < MAX_IN; ++d)
            if (d < D) gw1[d] = fmaf(xs[d], dh1, gw1[d]);
        __syncthreads();
    }
    float *g = partials + (int64_t)blockIdx.x * P;
#pragma unroll
    for (int d = 0; d < MAX_IN; ++d)
        if (d < D) g[o.w1 + d * H + j] = gw1[d];
    g[o.b1 + j] = gb1;
#pragma unroll
    for (int i = 0; i < H; ++i) g[o.w2 + i * H + j] = gw2[i];
    g[o.b2 + j] = gb2;
    g[o.w3 + j] = gw3;
    if (j == 0) {
        g[o.b3] = gb3;
        loss_partials[blockIdx.x] = loss;
    }
}

// grad[p] = sum over chunks c (in order) of partials[c][p]; loss likewise.
__global__ void mlp_reduce_kernel(const float *__restrict__ partials,
                                  const float *__restrict__ loss_partials, int64_t nchunks,
                                  int64_t P, float *__restrict__ grad, float *__restrict__ loss) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) {
        float acc = 0.0f;
        for (int64_t c = 0; c < nchunks; ++c) acc += partials[c * P + p];
        grad[p] = acc;
    }
    if (p == 0) {
        float l = 0.0f;
        for (int64_t c = 0; c < nchunks; ++c) l += loss_partials[c];
        *loss = l;
    }
}

__global__ void mlp_adam_kernel(float *__restrict__ theta, const float *__restrict__ grad,
                                float *__restrict__ m, float *__restrict__ v, int64_t P, float lr,
                                float beta1, float beta2, float eps, float bc1, float bc2,
                                float weight_decay) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float g = grad[p] + weight_decay * theta[p];
    const float mp = beta1 * m[p] + (1.0f - beta1) * g;
    const float vp = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
}

constexpr int64_t CHUNK = 32;

struct DeviceBuf {
    void *p = nullptr;
    ~DeviceBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int64_t mgx_mlp_param_count(int in_dim) {
    if (in_dim < 1 || in_dim > MAX_IN) return -1;
    return n_params(in_dim);
}

extern "C" int mgx_mlp_init(float *theta, int in_dim, uint64_t seed) {
    mgx::clear_error();
    MGX_ARG_CHECK(theta != nullptr, "theta is NULL");
    MGX_ARG_CHECK(in_dim >= 1 && in_dim <= MAX_IN, "in_dim out of range");
    // He-uniform weights from the splitmix64 stream, zero biases.
    const Offsets o = offsets(in_dim);
    const int64_t P = n_params(in_dim);
    uint64_t ctr = 0;
    auto uni = [&]() { return mgx::u64_to_uniform(mgx::splitmix64_at(seed, ++ctr)); };
    for (int64_t i = 0; i < P; ++i) theta[i] = 0.0f;
    const double a1 = std::sqrt(6.0 / in_dim), a2 = std::sqrt(6.0 / H), a3 = std::sqrt(3.0 / H);
    for (int i = 0; i < in_dim * H; ++i) theta[o.w1 + i] = (float)(a1 * uni());
    for (int i = 0; i < H * H; ++i) theta[o.w2 + i] = (float)(a2 * uni());
    for (int i = 0; i < H; ++i) theta[o.w3 + i] = (float)(a3 * uni());
    return MGX_OK;
}

// All arrays are host arrays; device buffers are managed internally.
extern "C" int mgx_mlp_predict(const float *theta, int in_dim, const float *X, int64_t n,
                               float *y) {
    mgx::clear_error();
    MGX_ARG_CHECK(theta && X && y, "NULL pointer");
    MGX_ARG_CHECK(in_dim >= 1 && in_dim <= MAX_IN && n >= 0, "invalid sizes");
    if (n == 0) return MGX_OK;
    const int64_t P = n_params(in_dim);
    DeviceBuf dth, dx, dy;
    MGX_HIP_TRY(hipMalloc(&dth.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dx.p, sizeof(float) * n * in_dim));
    MGX_HIP_TRY(hipMalloc(&dy.p, sizeof(float) * n));
    MGX_HIP_TRY(hipMemcpy(dth.p, theta, sizeof(float) * P, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dx.p, X, sizeof(float) * n * in_dim, hipMemcpyHostToDevice));
    const unsigned blocks = (unsigned)std::min<int64_t>(n, 4096);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(64), 0, 0, (const float *)dth.p,
                       (const float *)dx.p, n, in_dim, (float *)dy.p);
    int st = mgx::launch_status("mlp_forward_kernel");
    if (st != MGX_OK) return st;
    MGX_HIP_TRY(hipMemcpy(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MGX_OK;
}

// Mean-squared-error loss and its gradient at theta (host arrays).
extern "C" int mgx_mlp_gradient(const float *theta, int in_dim, const float *X, const float *t,
                                int64_t n, float *grad, float *loss) {
    mgx::clear_error();
    MGX_ARG_CHECK(theta && X && t && grad && loss, "NULL pointer");
    MGX_ARG_CHECK(in_dim >= 1 && in_dim <= MAX_IN && n > 0, "invalid sizes");
    const int64_t P = n_params(in_dim);
    const int64_t nchunks = (n + CHUNK - 1) / CHUNK;
    DeviceBuf dth, dx, dt, dpart, dlp, dg, dl;
    MGX_HIP_TRY(hipMalloc(&dth.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dx.p, sizeof(float) * n * in_dim));
    MGX_HIP_TRY(hipMalloc(&dt.p, sizeof(float) * n));
    MGX_HIP_TRY(hipMalloc(&dpart.p, sizeof(float) * P * nchunks));
    MGX_HIP_TRY(hipMalloc(&dlp.p, sizeof(float) * nchunks));
    MGX_HIP_TRY(hipMalloc(&dg.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dl.p, sizeof(float)));
    MGX_HIP_TRY(hipMemcpy(dth.p, theta, sizeof(float) * P, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dx.p, X, sizeof(float) * n * in_dim, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)nchunks), dim3(64), 0, 0,
                       (const float *)dth.p, (const float *)dx.p, (const float *)dt.p, n, in_dim,
                       CHUNK, 1.0f / (float)n, (float *)dpart.p, (float *)dlp.p);
    int st = mgx::launch_status("mlp_grad_kernel");
    if (st != MGX_OK) return st;
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, 0,
                       (const float *)dpart.p, (const float *)dlp.p, nchunks, P, (float *)dg.p,
                       (float *)dl.p);
    if ((st = mgx::launch_status("mlp_reduce_kernel")) != MGX_OK) return st;
    MGX_HIP_TRY(hipMemcpy(grad, dg.p, sizeof(float) * P, hipMemcpyDeviceToHost));
    MGX_HIP_TRY(hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    return MGX_OK;
}

// Full-batch Adam training for `epochs` steps, starting from theta (updated
// in place).  loss_history (may be NULL) receives the loss before each step.
extern "C" int mgx_mlp_train(float *theta, int in_dim, const float *X, const float *t, int64_t n,
                             int epochs, float lr, float weight_decay, float *loss_history) {
    mgx::clear_error();
    MGX_ARG_CHECK(theta && X && t, "NULL pointer");
    MGX_ARG_CHECK(in_dim >= 1 && in_dim <= MAX_IN && n > 0 && epochs >= 0, "invalid sizes");
    const int64_t P = n_params(in_dim);
    const int64_t nchunks = (n + CHUNK - 1) / CHUNK;
    DeviceBuf dth, dx, dt, dpart, dlp, dg, dl, dm, dv;
    MGX_HIP_TRY(hipMalloc(&dth.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dx.p, sizeof(float) * n * in_dim));
    MGX_HIP_TRY(hipMalloc(&dt.p, sizeof(float) * n));
    MGX_HIP_TRY(hipMalloc(&dpart.p, sizeof(float) * P * nchunks));
    MGX_HIP_TRY(hipMalloc(&dlp.p, sizeof(float) * nchunks));
    MGX_HIP_TRY(hipMalloc(&dg.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dl.p, sizeof(float) * (epochs > 0 ? epochs : 1)));
    MGX_HIP_TRY(hipMalloc(&dm.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMalloc(&dv.p, sizeof(float) * P));
    MGX_HIP_TRY(hipMemcpy(dth.p, theta, sizeof(float) * P, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dx.p, X, sizeof(float) * n * in_dim, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice));
    std::vector<float> zeros(P, 0.0f);
    MGX_HIP_TRY(hipMemcpy(dm.p, zeros.data(), sizeof(float) * P, hipMemcpyHostToDevice));
    MGX_HIP_TRY(hipMemcpy(dv.p, zeros.data(), sizeof(float) * P, hipMemcpyHostToDevice));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double p1 = 1.0, p2 = 1.0;
    int st;
    for (int e = 0; e < epochs; ++e) {
        hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)nchunks), dim3(64), 0, 0,
                           (const float *)dth.p, (const float *)dx.p, (const float *)dt.p, n,
                           in_dim, CHUNK, 1.0f / (float)n, (float *)dpart.p, (float *)dlp.p);
        if ((st = mgx::launch_status("mlp_grad_kernel")) != MGX_OK) return st;
        hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, 0,
                           (const float *)dpart.p, (const float *)dlp.p, nchunks, P,
                           (float *)dg.p, (float *)dl.p + e);
        if ((st = mgx::launch_status("mlp_reduce_kernel")) != MGX_OK) return st;
        p1 *= b1;
        p2 *= b2;
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, 0,
                           (float *)dth.p, (const float *)dg.p, (float *)dm.p, (float *)dv.p, P,
                           lr, b1, b2, eps, (float)(1.0 - p1), (float)(1.0 - p2), weight_decay);
        if ((st = mgx::launch_status("mlp_adam_kernel")) != MGX_OK) return st;
    }
    MGX_HIP_TRY(hipMemcpy(theta, dth.p, sizeof(float) * P, hipMemcpyDeviceToHost));
    if (loss_history && epochs > 0)
        MGX_HIP_TRY(hipMemcpy(loss_history, dl.p, sizeof(float) * epochs, hipMemcpyDeviceToHost));
    return MGX_OK;
}
