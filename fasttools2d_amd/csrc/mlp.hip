// The parameter selector's network: 8 inputs, two hidden layers of 64 GELU
// units, 4 outputs, FP32.  Forward pass, the gradient of the mean squared
// error, and Adam steps, all on the GPU.  No counterpart in the reference,
// whose drivers choose the Ewald parameters from analytic error estimates
// (StokesSLP_ewald_2p.m:84-104,147-202); those remain the fallback.
//
// Packed parameters (floats, in this order):
//   W1[64][8], b1[64], W2[64][64], b2[64], W3[4][64], b3[4]     (4996 in all)
// with z1 = W1 x + b1, a1 = gelu(z1), z2 = W2 a1 + b2, a2 = gelu(z2),
// y = W3 a2 + b3, and gelu(z) = z Phi(z) (the exact form, Phi the standard
// normal distribution function).
//
// Mapping: one 64-lane wave per tile of kTile samples; lane j owns hidden unit j
// of both layers (its weight rows in registers, W2 also in LDS for the
// backward pass's columns) and walks the tile's samples in order.  Gradients
// are summed per tile in registers, written as one partial per tile, and the
// partials are added in tile order by a second kernel, so every gradient is
// bit-reproducible.
#include "se2p_internal.h"

namespace se2p {

namespace {

constexpr int kIn = SE2P_MLP_IN, kHid = SE2P_MLP_HIDDEN, kOut = SE2P_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
              oW3 = oB2 + kHid, oB3 = oW3 + kOut * kHid, kNumParams = oB3 + kOut;
static_assert(kNumParams == SE2P_MLP_PARAMS, "packed parameter count");
static_assert(kHid == 64, "one lane per hidden unit of a 64-lane wave");
constexpr int kTile = 64;  // samples per wave

__device__ __forceinline__ float gelu(float z)
{
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

// d gelu / dz = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu_grad(float z)
{
    const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float phi = 0.39894228040143268f * __expf(-0.5f * z * z);
    return Phi + z * phi;
}

// Forward pass of samples [t*kTile, ...) (x: n x kIn, y: n x kOut).
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ w,
                                                    const float* __restrict__ x, int n,
                                                    float* __restrict__ y)
{
    __shared__ float a1s[kHid], a2s[kHid], xs[kIn];
    __shared__ float W2s[kHid][kHid + 1];
    const int j = threadIdx.x;
    for (int k = 0; k < kHid; ++k) W2s[j][k] = w[oW2 + j * kHid + k];
    float w1[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) w1[i] = w[oW1 + j * kIn + i];
    const float b1 = w[oB1 + j], b2 = w[oB2 + j];
    const int s0 = blockIdx.x * kTile, s1 = min(n, s0 + kTile);
    for (int sIdx = s0; sIdx < s1; ++sIdx) {
        __syncthreads();  // previous sample's activations have been read
        if (j < kIn) xs[j] = x[(size_t)sIdx * kIn + j];
        __syncthreads();
        float z = b1;
#pragma unroll
        for (int i = 0; i < kIn; ++i) z = fmaf(w1[i], xs[i], z);
        a1s[j] = gelu(z);
        __syncthreads();
        float z2 = b2;
        for (int k = 0; k < kHid; ++k) z2 = fmaf(W2s[j][k], a1s[k], z2);
        a2s[j] = gelu(z2);
        __syncthreads();
        if (j < kOut) {
            float o = w[oB3 + j];
            for (int k = 0; k < kHid; ++k) o = fmaf(w[oW3 + j * kHid + k], a2s[k], o);
            y[(size_t)sIdx * kOut + j] = o;
        }
    }
}

// Loss 1/(n kOut) sum (y - t)^2 over all samples and its gradient: tile t
// writes its partial gradient to part[t * kNumParams ...] and its partial
// loss to lpart[t].
__global__ __launch_bounds__(64) void k_mlp_grad_tiles(const float* __restrict__ w,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ tgt, int n,
                                                       float* __restrict__ part,
                                                       float* __restrict__ lpart)
{
    __shared__ float xs[kIn], a1s[kHid], a2s[kHid], d2s[kHid], dy[kOut];
    __shared__ float W2s[kHid][kHid + 1];
    __shared__ float W3s[kOut][kHid];
    const int j = threadIdx.x;
    for (int k = 0; k < kHid; ++k) W2s[j][k] = w[oW2 + j * kHid + k];
    for (int o = 0; o < kOut; ++o) W3s[o][j] = w[oW3 + o * kHid + j];
    float w1[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) w1[i] = w[oW1 + j * kIn + i];
    const float b1 = w[oB1 + j], b2 = w[oB2 + j];
    // gradient rows owned by lane j
    float gw1[kIn], gw2[kHid], gw3[kOut];
#pragma unroll
    for (int i = 0; i < kIn; ++i) gw1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < kHid; ++k) gw2[k] = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) gw3[o] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
    const float scale = 2.0f / ((float)n * kOut);
    const int s0 = blockIdx.x * kTile, s1 = min(n, s0 + kTile);
    for (int sIdx = s0; sIdx < s1; ++sIdx) {
        __syncthreads();
        if (j < kIn) xs[j] = x[(size_t)sIdx * kIn + j];
        __syncthreads();
        float z1 = b1;
#pragma unroll
        for (int i = 0; i < kIn; ++i) z1 = fmaf(w1[i], xs[i], z1);
        a1s[j] = gelu(z1);
        __syncthreads();
        float z2 = b2;
#pragma unroll
        for (int k = 0; k < kHid; ++k) z2 = fmaf(W2s[j][k], a1s[k], z2);
        a2s[j] = gelu(z2);
        __syncthreads();
        if (j < kOut) {
            float o = w[oB3 + j];
            for (int k = 0; k < kHid; ++k) o = fmaf(W3s[j][k], a2s[k], o);
            const float r = o - tgt[(size_t)sIdx * kOut + j];
            loss = fmaf(r, r, loss);
            dy[j] = scale * r;
        }
        __syncthreads();
        // output layer: lane j owns column j of W3
        float back = 0.f;
#pragma unroll
        for (int o = 0; o < kOut; ++o) {
            gw3[o] = fmaf(dy[o], a2s[j], gw3[o]);
            back = fmaf(W3s[o][j], dy[o], back);
        }
        if (j < kOut) gb3 += dy[j];
        const float d2 = back * gelu_grad(z2);
        d2s[j] = d2;
        gb2 += d2;
#pragma unroll
        for (int k = 0; k < kHid; ++k) gw2[k] = fmaf(d2, a1s[k], gw2[k]);
        __syncthreads();
        // first layer: lane j owns hidden unit j, column j of W2
        float b = 0.f;
#pragma unroll
        for (int k = 0; k < kHid; ++k) b = fmaf(W2s[k][j], d2s[k], b);
        const float d1 = b * gelu_grad(z1);
        gb1 += d1;
#pragma unroll
        for (int i = 0; i < kIn; ++i) gw1[i] = fmaf(d1, xs[i], gw1[i]);
    }
    float* g = part + (size_t)blockIdx.x * kNumParams;
#pragma unroll
    for (int i = 0; i < kIn; ++i) g[oW1 + j * kIn + i] = gw1[i];
    g[oB1 + j] = gb1;
#pragma unroll
    for (int k = 0; k < kHid; ++k) g[oW2 + j * kHid + k] = gw2[k];
    g[oB2 + j] = gb2;
#pragma unroll
    for (int o = 0; o < kOut; ++o) g[oW3 + o * kHid + j] = gw3[o];
    if (j < kOut) g[oB3 + j] = gb3;
    // the tile's loss: lanes 0..kOut-1 hold it
    __shared__ float ls[kOut];
    if (j < kOut) ls[j] = loss;
    __syncthreads();
    if (j == 0) {
        float s = 0.f;
        for (int o = 0; o < kOut; ++o) s += ls[o];
        lpart[blockIdx.x] = s / ((float)n * kOut);
    }
}

// grad[p] = sum over tiles of part[t][p], in tile order; loss likewise.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ lpart,
                             int ntiles, float* __restrict__ grad, float* __restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNumParams) {
        float s = 0.f;
        for (int t = 0; t < ntiles; ++t) s += part[(size_t)t * kNumParams + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.f;
        for (int t = 0; t < ntiles; ++t) s += lpart[t];
        *loss = s;
    }
}

// Adam (Kingma & Ba, 2015) with bias correction, step = 1, 2, ...
__global__ void k_mlp_adam(float* __restrict__ w, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, int step, float lr,
                           float b1, float b2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float mh = mp / (1.f - powf(b1, (float)step));
    const float vh = vp / (1.f - powf(b2, (float)step));
    w[p] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

int mlp_forward(se2p_ctx* ctx, const float* w, const float* x, int n, float* y)
{
    if (n <= 0) return SE2P_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + kTile - 1) / kTile), dim3(64), 0, ctx->stream, w,
                       x, n, y);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int mlp_grad(se2p_ctx* ctx, Workspace* ws, const float* w, const float* x, const float* t, int n,
             float* grad, float* loss)
{
    if (n <= 0) return set_error(SE2P_ERR_ARG, "mlp: no samples");
    const int ntiles = (n + kTile - 1) / kTile;
    void* part;
    SE2P_TRY(ws_get(ctx, ws, WS_MLP_PART,
                    sizeof(float) * ((size_t)ntiles * kNumParams + (size_t)ntiles), &part));
    float* pp = (float*)part;
    float* lp = pp + (size_t)ntiles * kNumParams;
    hipLaunchKernelGGL(k_mlp_grad_tiles, dim3(ntiles), dim3(64), 0, ctx->stream, w, x, t, n, pp,
                       lp);
    SE2P_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNumParams + 255) / 256), dim3(256), 0, ctx->stream, pp,
                       lp, ntiles, grad, loss);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

int mlp_adam(se2p_ctx* ctx, float* w, const float* g, float* m, float* v, int step, float lr,
             float b1, float b2, float eps)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, ctx->stream, w, g,
                       m, v, step, lr, b1, b2, eps);
    SE2P_HIP_CHECK(hipGetLastError());
    return SE2P_OK;
}

}  // namespace se2p
