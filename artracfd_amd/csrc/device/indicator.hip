// indicator.hip - training and batch inference of the learned troubled-face
// indicator on BF16 MFMA (FP32 accumulation), and face sampling from the
// solver's own fields.
//
// Parameter vector (FP32 master copy, kMlpParams values):
//   [0, 512)      W1[16][32]  (rows 12..15 stay zero)
//   [512, 544)    b1[32]
//   [544, 1568)   W2[32][32]
//   [1568, 1600)  b2[32]
//   [1600, 1632)  w3[32]
//   [1632]        b3
// Training minimises the weighted logistic loss of "troubled" (label > tau)
// with Adam.  One block (one wavefront) handles 64 samples: forward with
// MFMA, backward with MFMA for dW2 = H1^T dZ2, dH1 = dZ2 W2^T and
// dW1 = X^T dZ1; per-block partial gradients are summed by a second kernel
// in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>

#include "indicator.hpp"

extern "C" {
#include "artra_device.h"
}

using namespace artra;

namespace {

constexpr int kParams = ARTRA_MLP_PARAMS;
constexpr int oW1 = 0, oB1 = 512, oW2 = 544, oB2 = 1568, oW3 = 1600, oB3 = 1632;

thread_local char gIndError[256];

int Fail(const char *fmt, ...)
{
    va_list a;
    va_start(a, fmt);
    vsnprintf(gIndError, sizeof gIndError, fmt, a);
    va_end(a);
    return 1;
}

#define HIP_OK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return Fail("%s failed: %s", #expr, hipGetErrorString(e_));                 \
    } while (0)

__device__ __forceinline__ void Sync() { __syncthreads(); }

__device__ __forceinline__ floatx16 Zero16()
{
    floatx16 z = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    return z;
}

// bf16 copies of W1 and W2 from the FP32 master parameters
__global__ void mlp_pack(const float *params, uint16_t *w1, uint16_t *w2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 512) w1[t] = f32_to_bf16(params[oW1 + t]);
    if (t < 1024) w2[t] = f32_to_bf16(params[oW2 + t]);
}

struct TrainScratch {
    uint16_t X[64 * kFeatPad];
    float H1[64 * kHidden];
    float H2[64 * kHidden];
    float dZ[64 * kHidden];
    float dlogit[64];
};

// Forward and backward over one 64-sample chunk (one wavefront per block).
// partial[block][kParams] receives this chunk's gradient; loss[block] its
// weighted loss sum.  Samples past n contribute nothing.
__global__ __launch_bounds__(64) void mlp_grad(const float *params, const uint16_t *w1b, const uint16_t *w2b,
                                               const float *feat, const float *label, long n, float tau,
                                               float wpos, float inv_weight_sum, float *partial, float *loss)
{
    __shared__ TrainScratch sc;
    const int l = threadIdx.x;
    const int j = l & 31, half = l >> 5;
    const long s = (long)blockIdx.x * 64 + l;
    const bool valid = s < n;
    MlpDev net{w1b, w2b, params + oB1, params + oB2, params + oW3, params[oB3], 0.0f};
    MlpFrags f;
    mlp_fragments(net, f);
#pragma unroll
    for (int k = 0; k < kFeatPad; ++k) {
        sc.X[l * kFeatPad + k] = (valid && k < kFeat) ? f32_to_bf16(feat[s * kFeat + k]) : (uint16_t)0;
    }
    Sync();
    // forward: H1 = relu(X W1 + b1), H2 = relu(bf16(H1) W2 + b2), z = H2 . w3 + b3
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        floatx16 acc = Zero16();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x4 a;
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = (short)sc.X[(32 * rb + j) * kFeatPad + 8 * ks + 4 * half + t];
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, f.w1[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            sc.H1[(32 * rb + row) * kHidden + j] = bf16_to_f32(f32_to_bf16(fmaxf(acc[r] + f.b1, 0.0f)));
        }
    }
    Sync();
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        floatx16 acc = Zero16();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 a;
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = (short)f32_to_bf16(sc.H1[(32 * rb + j) * kHidden + 8 * ks + 4 * half + t]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, f.w2[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            sc.H2[(32 * rb + row) * kHidden + j] = fmaxf(acc[r] + f.b2, 0.0f);
        }
    }
    Sync();
    float z = params[oB3];
    for (int k = 0; k < kHidden; ++k) z += sc.H2[l * kHidden + k] * params[oW3 + k];
    // weighted logistic loss
    float dl = 0.0f, ls = 0.0f;
    if (valid) {
        const float y = (label[s] > tau) ? 1.0f : 0.0f;
        const float w = (y > 0.5f ? wpos : 1.0f) * inv_weight_sum;
        const float sig = 1.0f / (1.0f + expf(-z));
        dl = (sig - y) * w;
        ls = w * (fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z))));
    }
    sc.dlogit[l] = dl;
    Sync();
    // dZ2 = dlogit w3 (H2 > 0)
    for (int k = 0; k < kHidden; ++k) {
        sc.dZ[l * kHidden + k] = (sc.H2[l * kHidden + k] > 0.0f) ? dl * params[oW3 + k] : 0.0f;
    }
    Sync();
    float *g = partial + (long)blockIdx.x * kParams;
    // w3, b3, b2 by lane-owned columns (fixed order over the 64 samples)
    if (l < 32) {
        float a3 = 0.0f, a2 = 0.0f;
        for (int q = 0; q < 64; ++q) {
            a3 += sc.H2[q * kHidden + l] * sc.dlogit[q];
            a2 += sc.dZ[q * kHidden + l];
        }
        g[oW3 + l] = a3;
        g[oB2 + l] = a2;
    }
    if (l == 0) {
        float a = 0.0f;
        for (int q = 0; q < 64; ++q) a += sc.dlogit[q];
        g[oB3] = a;
    }
    // dW2[k][jj] = sum_q H1[q][k] dZ2[q][jj]: A = H1^T (32 x 64), B = dZ2 (64 x 32)
    {
        floatx16 acc = Zero16();
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = 8 * ks + 4 * half + t;
                a[t] = (short)f32_to_bf16(sc.H1[q * kHidden + j]);
                b[t] = (short)f32_to_bf16(sc.dZ[q * kHidden + j]);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            g[oW2 + row * kHidden + j] = acc[r];
        }
    }
    // dH1[q][k] = sum_jj dZ2[q][jj] W2[k][jj]: A = dZ2 (64 x 32), B = W2^T (32 x 32)
    bf16x4 w2t[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int t = 0; t < 4; ++t) w2t[ks][t] = (short)w2b[j * kHidden + 8 * ks + 4 * half + t];
    float dH1[2][16];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        floatx16 acc = Zero16();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 a;
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = (short)f32_to_bf16(sc.dZ[(32 * rb + j) * kHidden + 8 * ks + 4 * half + t]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, w2t[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) dH1[rb][r] = acc[r];
    }
    Sync();
    // dZ1 = dH1 (H1 > 0), stored over dZ
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * rb + 8 * (r >> 2) + 4 * half + (r & 3);
            sc.dZ[row * kHidden + j] = (sc.H1[row * kHidden + j] > 0.0f) ? dH1[rb][r] : 0.0f;
        }
    Sync();
    if (l < 32) {
        float a1 = 0.0f;
        for (int q = 0; q < 64; ++q) a1 += sc.dZ[q * kHidden + l];
        g[oB1 + l] = a1;
    }
    // dW1[i][k] = sum_q X[q][i] dZ1[q][k]: A = X^T (rows i < 16, padded to 32), B = dZ1
    {
        floatx16 acc = Zero16();
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            bf16x4 a, b;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = 8 * ks + 4 * half + t;
                a[t] = (j < kFeatPad) ? (short)sc.X[q * kFeatPad + j] : (short)0;
                b[t] = (short)f32_to_bf16(sc.dZ[q * kHidden + j]);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            if (row < kFeatPad) g[oW1 + row * kHidden + j] = (row < kFeat) ? acc[r] : 0.0f;
        }
    }
    // per-block loss
    sc.dlogit[l] = ls;
    Sync();
    if (l == 0) {
        float a = 0.0f;
        for (int q = 0; q < 64; ++q) a += sc.dlogit[q];
        loss[blockIdx.x] = a;
    }
}

// Sum the per-block partial gradients in block order.
__global__ void mlp_reduce(const float *partial, int blocks, float *grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float a = 0.0f;
    for (int b = 0; b < blocks; ++b) a += partial[(long)b * kParams + p];
    grad[p] = a;
}

__global__ void mlp_adam(float *params, float *m, float *v, const float *grad, float lr, float b1, float b2,
                         float eps, float corr1, float corr2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    if (p >= oW1 + 12 * kHidden && p < oB1) return;   // padded input rows stay zero
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / corr1) / (sqrtf(vv / corr2) + eps);
}

// Batch inference (64 samples per block) through the same wave routine the
// sweeps use.
__global__ __launch_bounds__(64) void mlp_infer(const float *params, const uint16_t *w1b, const uint16_t *w2b,
                                                const float *feat, long n, float *logit)
{
    __shared__ float bias[3 * kHidden];
    const int l = threadIdx.x;
    if (l < kHidden) {
        bias[l] = params[oB1 + l];
        bias[kHidden + l] = params[oB2 + l];
        bias[2 * kHidden + l] = params[oW3 + l];
    }
    Sync();
    const long s = (long)blockIdx.x * 64 + l;
    float x[kFeat];
#pragma unroll
    for (int k = 0; k < kFeat; ++k) x[k] = (s < n) ? feat[s * kFeat + k] : 0.0f;
    MlpWave f;
    mlp_wave_fragments(w1b, w2b, f);
    const float z = mlp_logit_wave(x, f, bias, params[oB3]);
    if (s < n) logit[s] = z;
}

struct Trainer {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *partial = nullptr, *loss = nullptr;
    uint16_t *w1 = nullptr, *w2 = nullptr;
    float *feat = nullptr, *label = nullptr;
    long n = 0;
    int maxBlocks = 0;
    ~Trainer()
    {
        hipFree(params); hipFree(m); hipFree(v); hipFree(grad); hipFree(partial); hipFree(loss);
        hipFree(w1); hipFree(w2); hipFree(feat); hipFree(label);
    }
};

int Setup(Trainer &t, const float *params, const float *feat, const float *label, long n, int batch)
{
    t.n = n;
    t.maxBlocks = (batch + 63) / 64;
    HIP_OK(hipMalloc(&t.params, kParams * sizeof(float)));
    HIP_OK(hipMalloc(&t.m, kParams * sizeof(float)));
    HIP_OK(hipMalloc(&t.v, kParams * sizeof(float)));
    HIP_OK(hipMalloc(&t.grad, kParams * sizeof(float)));
    HIP_OK(hipMalloc(&t.partial, (size_t)t.maxBlocks * kParams * sizeof(float)));
    HIP_OK(hipMalloc(&t.loss, (size_t)t.maxBlocks * sizeof(float)));
    HIP_OK(hipMalloc(&t.w1, 512 * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&t.w2, 1024 * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&t.feat, (size_t)n * kFeat * sizeof(float)));
    if (label) HIP_OK(hipMalloc(&t.label, (size_t)n * sizeof(float)));
    HIP_OK(hipMemcpy(t.params, params, kParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(t.m, 0, kParams * sizeof(float)));
    HIP_OK(hipMemset(t.v, 0, kParams * sizeof(float)));
    HIP_OK(hipMemcpy(t.feat, feat, (size_t)n * kFeat * sizeof(float), hipMemcpyHostToDevice));
    if (label) HIP_OK(hipMemcpy(t.label, label, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace

extern "C" {

const char *artra_mlp_error(void)
{
    return gIndError;
}

int artra_mlp_gradient_dev(const float *params, const float *feat, const float *label, long n, float tau, float wpos,
                           float *grad, float *loss)
{
    Trainer t;
    if (Setup(t, params, feat, label, n, (int)n)) return 1;
    double wsum = 0.0;
    for (long q = 0; q < n; ++q) wsum += (label[q] > tau) ? wpos : 1.0;
    const int blocks = (int)((n + 63) / 64);
    hipLaunchKernelGGL(mlp_pack, dim3(4), dim3(256), 0, 0, t.params, t.w1, t.w2);
    hipLaunchKernelGGL(mlp_grad, dim3(blocks), dim3(64), 0, 0, t.params, t.w1, t.w2, t.feat, t.label, n, tau, wpos,
                       (float)(1.0 / wsum), t.partial, t.loss);
    hipLaunchKernelGGL(mlp_reduce, dim3((kParams + 255) / 256), dim3(256), 0, 0, t.partial, blocks, t.grad);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(grad, t.grad, kParams * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> ls(blocks);
    HIP_OK(hipMemcpy(ls.data(), t.loss, blocks * sizeof(float), hipMemcpyDeviceToHost));
    double a = 0.0;
    for (float x : ls) a += x;
    *loss = (float)a;
    return 0;
}

int artra_mlp_train_dev(float *params, const float *feat, const float *label, long n, const ArtraMlpTraining *opt,
                        float *loss_history)
{
    const int batch = opt->batch > 0 ? opt->batch : 4096;
    if (n < batch) return Fail("training needs at least one batch (%ld samples < batch %d)", n, batch);
    if (opt->epochs < 1) return Fail("training needs at least one epoch");
    Trainer t;
    if (Setup(t, params, feat, label, n, batch)) return 1;
    // class weights: positives weighted by wpos, normalised per mini-batch by its expected weight
    long pos = 0;
    for (long q = 0; q < n; ++q) pos += (label[q] > opt->tau);
    const double frac = (double)pos / (double)(n > 0 ? n : 1);
    const double wbatch = batch * (frac * opt->wpos + (1.0 - frac));
    hipLaunchKernelGGL(mlp_pack, dim3(4), dim3(256), 0, 0, t.params, t.w1, t.w2);
    const long batches = n / batch;
    long step = 0;
    for (int ep = 0; ep < opt->epochs; ++ep) {
        for (long bq = 0; bq < batches; ++bq, ++step) {
            // permutation-free schedule: batches visit the (pre-shuffled) data in a fixed rotated order
            const long first = ((bq + (long)ep * 7919) % batches) * batch;
            const int blocks = (batch + 63) / 64;
            const double progress = (double)step / (double)(batches * opt->epochs);
            const float lr = (float)(opt->lr * 0.5 * (1.0 + cos(3.141592653589793 * progress)));
            hipLaunchKernelGGL(mlp_grad, dim3(blocks), dim3(64), 0, 0, t.params, t.w1, t.w2, t.feat + first * kFeat,
                               t.label + first, (long)batch, opt->tau, opt->wpos, (float)(1.0 / wbatch), t.partial,
                               t.loss);
            hipLaunchKernelGGL(mlp_reduce, dim3((kParams + 255) / 256), dim3(256), 0, 0, t.partial, blocks, t.grad);
            const float c1 = 1.0f - powf(0.9f, (float)(step + 1)), c2 = 1.0f - powf(0.999f, (float)(step + 1));
            hipLaunchKernelGGL(mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, 0, t.params, t.m, t.v, t.grad, lr,
                               0.9f, 0.999f, 1e-8f, c1, c2);
            hipLaunchKernelGGL(mlp_pack, dim3(4), dim3(256), 0, 0, t.params, t.w1, t.w2);
            if (loss_history && bq == batches - 1) {
                std::vector<float> ls(blocks);
                HIP_OK(hipMemcpy(ls.data(), t.loss, blocks * sizeof(float), hipMemcpyDeviceToHost));
                double a = 0.0;
                for (float x : ls) a += x;
                loss_history[ep] = (float)a;
            }
        }
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(params, t.params, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int artra_mlp_infer_dev(const float *params, const float *feat, long n, float *logit)
{
    Trainer t;
    if (Setup(t, params, feat, nullptr, n, 64)) return 1;
    float *out = nullptr;
    HIP_OK(hipMalloc(&out, (size_t)n * sizeof(float)));
    hipLaunchKernelGGL(mlp_pack, dim3(4), dim3(256), 0, 0, t.params, t.w1, t.w2);
    hipLaunchKernelGGL(mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, t.params, t.w1, t.w2, t.feat, n,
                       out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(logit, out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(out);
    if (e != hipSuccess) return Fail("inference failed: %s", hipGetErrorString(e));
    return 0;
}

}  // extern "C"
