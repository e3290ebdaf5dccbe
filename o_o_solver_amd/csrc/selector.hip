// The learned precision selector: a small MLP (8 → 32 → 32 → 1, tanh hidden layers, a logistic output) that gives
// p(FP32 fronts + FP64 refinement reach the target backward error) from features of the structure and the matrix.
// FP32 throughout; forward, backward and Adam steps as HIP kernels, every reduction in a fixed order (deterministic).
//
// Parameters, one flat array (PARAMETERS floats): W1 (32 × 8, row-major: W1[o·8 + i]), b1 (32), W2 (32 × 32), b2
// (32), W3 (32), b3 (1). Batches are row-major count × 8 features.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "selector.h"

namespace oosg
{

namespace
{

constexpr int F = SELECTOR_FEATURES, H = SELECTOR_HIDDEN;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * F, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H;
static_assert(OFF_B3 + 1 == SELECTOR_PARAMETERS, "parameter layout");

// One thread per sample: z1 = W1·x + b1, a1 = tanh z1, z2 = W2·a1 + b2, a2 = tanh z2, out = W3·a2 + b3,
// p = 1 / (1 + e^−out). a1, a2 (count × H) are kept for the backward pass when not null.
__global__ __launch_bounds__(256) void selectorForwardKernel(const float* params, const float* x, int64_t count, float* a1,
                                                     float* a2, float* probability, float* logit)
{
    __shared__ float w[SELECTOR_PARAMETERS];
    for (int q = threadIdx.x; q < SELECTOR_PARAMETERS; q += blockDim.x)
        w[q] = params[q];
    __syncthreads();
    const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (b >= count)
        return;
    float in[F];
    for (int i = 0; i < F; ++i)
        in[i] = x[b * F + i];
    float h1[H];
    for (int o = 0; o < H; ++o)
    {
        float z = w[OFF_B1 + o];
        for (int i = 0; i < F; ++i)
            z += w[OFF_W1 + o * F + i] * in[i];
        h1[o] = tanhf(z);
        if (a1)
            a1[b * H + o] = h1[o];
    }
    float out = w[OFF_B3];
    for (int o = 0; o < H; ++o)
    {
        float z = w[OFF_B2 + o];
        for (int i = 0; i < H; ++i)
            z += w[OFF_W2 + o * H + i] * h1[i];
        const float h2 = tanhf(z);
        if (a2)
            a2[b * H + o] = h2;
        out += w[OFF_W3 + o] * h2;
    }
    if (logit)
        logit[b] = out;
    probability[b] = 1.0f / (1.0f + expf(-out));
}

// Per sample the output error d3 = (p − y)/count (the mean binary cross-entropy's derivative by the logit) and the
// back-propagated d2 = (W3 ∘ d3)·(1 − a2²), d1 = (W2ᵀ·d2)·(1 − a1²).
__global__ __launch_bounds__(256) void selectorDeltaKernel(const float* params, const float* probability, const float* labels,
                                                   const float* a1, const float* a2, int64_t count, float* d3,
                                                   float* d2, float* d1)
{
    __shared__ float w[SELECTOR_PARAMETERS];
    for (int q = threadIdx.x; q < SELECTOR_PARAMETERS; q += blockDim.x)
        w[q] = params[q];
    __syncthreads();
    const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (b >= count)
        return;
    const float e = (probability[b] - labels[b]) / float(count);
    d3[b] = e;
    float g2[H];
    for (int o = 0; o < H; ++o)
    {
        const float h = a2[b * H + o];
        g2[o] = w[OFF_W3 + o] * e * (1.0f - h * h);
        d2[b * H + o] = g2[o];
    }
    for (int i = 0; i < H; ++i)
    {
        float s = 0.0f;
        for (int o = 0; o < H; ++o)
            s += w[OFF_W2 + o * H + i] * g2[o];
        const float h = a1[b * H + i];
        d1[b * H + i] = s * (1.0f - h * h);
    }
}

// One thread per parameter: its gradient Σ_b (delta of its output)·(input of its weight), over the samples in order.
__global__ __launch_bounds__(256) void selectorGradientKernel(const float* x, const float* a1, const float* a2,
                                                      const float* d3, const float* d2, const float* d1,
                                                      int64_t count, float* gradient)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= SELECTOR_PARAMETERS)
        return;
    // (delta array, its width, its column; input array, its width, its column; -1: a bias)
    const float* delta;
    int dw, dc;
    const float* input = nullptr;
    int iw = 0, ic = 0;
    if (q < OFF_B1)
    {
        delta = d1, dw = H, dc = (q - OFF_W1) / F;
        input = x, iw = F, ic = (q - OFF_W1) % F;
    }
    else if (q < OFF_W2)
        delta = d1, dw = H, dc = q - OFF_B1;
    else if (q < OFF_B2)
    {
        delta = d2, dw = H, dc = (q - OFF_W2) / H;
        input = a1, iw = H, ic = (q - OFF_W2) % H;
    }
    else if (q < OFF_W3)
        delta = d2, dw = H, dc = q - OFF_B2;
    else if (q < OFF_B3)
    {
        delta = d3, dw = 1, dc = 0;
        input = a2, iw = H, ic = q - OFF_W3;
    }
    else
        delta = d3, dw = 1, dc = 0;
    float s = 0.0f;
    for (int64_t b = 0; b < count; ++b)
        s += delta[b * dw + dc] * (input ? input[b * iw + ic] : 1.0f);
    gradient[q] = s;
}

// The mean binary cross-entropy, by one workgroup in a fixed order: −mean(y·log p + (1 − y)·log(1 − p)), from the
// logits (log p = −softplus(−z), log(1 − p) = −softplus(z)).
__global__ __launch_bounds__(256) void selectorLossKernel(const float* logit, const float* labels, int64_t count, float* loss)
{
    __shared__ float part[256];
    float s = 0.0f;
    for (int64_t b = threadIdx.x; b < count; b += blockDim.x)
    {
        const float z = logit[b], y = labels[b];
        const float softMinus = fmaxf(-z, 0.0f) + log1pf(expf(-fabsf(z))); // softplus(−z)
        const float softPlus = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));   // softplus(z)
        s += y * softMinus + (1.0f - y) * softPlus;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float total = 0.0f;
        for (int t = 0; t < 256; ++t)
            total += part[t];
        loss[0] = total / float(count);
    }
}

// Adam (Kingma & Ba 2015): m = β1 m + (1 − β1) g, v = β2 v + (1 − β2) g², θ −= lr·m̂/(sqrt(v̂) + ε).
__global__ void selectorAdamKernel(float* params, const float* gradient, float* m, float* v, float lr, float correction1,
                           float correction2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= SELECTOR_PARAMETERS)
        return;
    const float g = gradient[q];
    m[q] = 0.9f * m[q] + 0.1f * g;
    v[q] = 0.999f * v[q] + 0.001f * g * g;
    const float mh = m[q] / correction1, vh = v[q] / correction2;
    params[q] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

int groupsFor(int64_t n)
{
    return int((n + 255) / 256);
}

} // namespace

hipError_t selectorForward(const float* params, const float* x, int64_t count, float* a1, float* a2,
                           float* probability, float* logit, hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(selectorForwardKernel, dim3(groupsFor(count)), dim3(256), 0, stream, params, x, count, a1, a2,
                       probability, logit);
    return hipGetLastError();
}

hipError_t selectorGradient(const float* params, const float* x, const float* labels, int64_t count, float* a1,
                            float* a2, float* probability, float* logit, float* d3, float* d2, float* d1,
                            float* gradient, float* loss, hipStream_t stream)
{
    if (count <= 0)
        return hipErrorInvalidValue;
    hipError_t e = selectorForward(params, x, count, a1, a2, probability, logit, stream);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(selectorDeltaKernel, dim3(groupsFor(count)), dim3(256), 0, stream, params, probability, labels, a1, a2,
                       count, d3, d2, d1);
    hipLaunchKernelGGL(selectorGradientKernel, dim3(groupsFor(SELECTOR_PARAMETERS)), dim3(256), 0, stream, x, a1, a2, d3, d2,
                       d1, count, gradient);
    hipLaunchKernelGGL(selectorLossKernel, dim3(1), dim3(256), 0, stream, logit, labels, count, loss);
    return hipGetLastError();
}

hipError_t selectorAdam(float* params, const float* gradient, float* m, float* v, float lr, int step,
                        hipStream_t stream)
{
    const float c1 = 1.0f - powf(0.9f, float(step)), c2 = 1.0f - powf(0.999f, float(step));
    hipLaunchKernelGGL(selectorAdamKernel, dim3(groupsFor(SELECTOR_PARAMETERS)), dim3(256), 0, stream, params, gradient, m, v,
                       lr, c1, c2);
    return hipGetLastError();
}

} // namespace oosg
