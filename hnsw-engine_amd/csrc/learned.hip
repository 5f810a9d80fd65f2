// The learned stop: a 16 -> 64 -> 1 MLP (bfloat16 weights, float32 accumulation) that predicts,
// after each layer-0 expansion of the register-beam search, whether the current top-k is already
// final. Training runs on the GPU with these kernels: a forward/backward pass over a batch
// (binary cross-entropy on the logit, per-block partial gradients summed in a fixed order, so
// training is deterministic) and an Adam step on the float32 master weights. The weights the
// kernels use are the master weights rounded to bfloat16 (round to nearest even).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/hnsw_mi355x.h"
#include "common.hpp"
#include "index.hpp"
#include "search.hpp"
#include "wave.hpp"

namespace hm {

constexpr int kGradBlocks = 128;

HM_DEV float bf16_round(float x) {
    uint32_t b = __float_as_uint(x);
    b += 0x7fffu + ((b >> 16) & 1u);
    return __uint_as_float(b & 0xffff0000u);
}

// Forward and backward over samples [0, n) (rows of x gathered through perm when set).
// Block b writes its partial gradient (kModelSize floats; the normalisation slots stay 0) and its
// partial loss to part[b]. Lane j of every wave owns hidden unit j.
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *x, const float *y,
                                                       const int32_t *perm, int64_t n,
                                                       const float *w, float pos_weight,
                                                       float *part) {
    __shared__ float red[4][kModelSize + 1];
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    float w1[kFeatures], mean[kFeatures], istd[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) {
        w1[i] = bf16_round(w[kModelW1 + i * kHidden + lane]);
        mean[i] = w[i];
        istd[i] = w[kFeatures + i];
    }
    const float b1 = bf16_round(w[kModelB1 + lane]);
    const float w2 = bf16_round(w[kModelW2 + lane]);
    const float b2 = bf16_round(w[kModelB2]);
    float g1[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) g1[i] = 0.f;
    float gb1 = 0.f, gw2 = 0.f, gb2 = 0.f, loss = 0.f;
    const int64_t stride = (int64_t)gridDim.x * 4;
    for (int64_t s = (int64_t)blockIdx.x * 4 + wave; s < n; s += stride) {
        const int64_t row = perm ? perm[s] : s;
        float xn[kFeatures];
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) xn[i] = (x[row * kFeatures + i] - mean[i]) * istd[i];
        float z = b1;
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) z += w1[i] * xn[i];
        const float h = fmaxf(z, 0.f);
        float o = w2 * h;
        for (int m = 32; m >= 1; m >>= 1) o += __shfl_xor(o, m, 64);
        o += b2;
        const float t = y[row];
        const float sp_pos = o > 0.f ? log1pf(expf(-o)) : -o + log1pf(expf(o));  // softplus(-o)
        const float sp_neg = o + sp_pos;                                           // softplus(o)
        const float sig = 1.f / (1.f + expf(-o));
        loss += pos_weight * t * sp_pos + (1.f - t) * sp_neg;
        const float dout = pos_weight * t * (sig - 1.f) + (1.f - t) * sig;
        gb2 += dout;
        gw2 += dout * h;
        const float dh = z > 0.f ? dout * w2 : 0.f;
        gb1 += dh;
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) g1[i] += dh * xn[i];
    }
    // Block reduction in a fixed order: each wave writes its partials, wave 0 adds them.
    float *r = red[wave];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) r[kModelW1 + i * kHidden + lane] = g1[i];
    r[kModelB1 + lane] = gb1;
    r[kModelW2 + lane] = gw2;
    if (lane == 0) {
        r[kModelB2] = gb2;
        r[kModelSize] = loss;
    }
    __syncthreads();
    float *out = part + (size_t)blockIdx.x * (kModelSize + 1);
    for (int i = threadIdx.x; i <= kModelSize; i += blockDim.x) {
        const float v = i < kModelW1 ? 0.f : red[0][i] + red[1][i] + red[2][i] + red[3][i];
        out[i] = v;
    }
}

// Sums the block partials in block order and scales by 1/n.
__global__ void mlp_reduce_kernel(const float *part, int blocks, int64_t n, float *grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kModelSize) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += part[(size_t)b * (kModelSize + 1) + i];
    grad[i] = s / (float)n;
}

__global__ void adam_kernel(float *w, const float *grad, float *m, float *v, float lr, float b1,
                            float b2, float eps, int step) {
    const int i = kModelW1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kModelSize) return;
    const float g = grad[i];
    m[i] = b1 * m[i] + (1.f - b1) * g;
    v[i] = b2 * v[i] + (1.f - b2) * g * g;
    const float mh = m[i] / (1.f - powf(b1, (float)step));
    const float vh = v[i] / (1.f - powf(b2, (float)step));
    w[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *x, int64_t n, const float *w,
                                                          float *logit) {
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    float w1[kFeatures], mean[kFeatures], istd[kFeatures];
#pragma unroll
    for (int i = 0; i < kFeatures; ++i) {
        w1[i] = bf16_round(w[kModelW1 + i * kHidden + lane]);
        mean[i] = w[i];
        istd[i] = w[kFeatures + i];
    }
    const float b1 = bf16_round(w[kModelB1 + lane]);
    const float w2 = bf16_round(w[kModelW2 + lane]);
    const float b2 = bf16_round(w[kModelB2]);
    for (int64_t s = (int64_t)blockIdx.x * 4 + wave; s < n; s += (int64_t)gridDim.x * 4) {
        float z = b1;
#pragma unroll
        for (int i = 0; i < kFeatures; ++i) z += w1[i] * ((x[s * kFeatures + i] - mean[i]) * istd[i]);
        float o = w2 * fmaxf(z, 0.f);
        for (int m = 32; m >= 1; m >>= 1) o += __shfl_xor(o, m, 64);
        if (lane == 0) logit[s] = o + b2;
    }
}

}  // namespace hm

using namespace hm;

namespace {

float bf16_host(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    b += 0x7fffu + ((b >> 16) & 1u);
    b &= 0xffff0000u;
    memcpy(&x, &b, 4);
    return x;
}

template <class T>
struct Dev {
    T *p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

int hm_model_size(void) { return kModelSize; }

int hm_mlp_grad(const float *x, const float *y, int64_t n, const float *params, float pos_weight,
                float *grad, float *loss) {
    if (!x || !y || !params || !grad || n < 1) return invalid("bad arguments");
    Dev<float> dx, dy, dw, part, dg;
    HM_TRY(dx.alloc((size_t)n * kFeatures), "alloc");
    HM_TRY(dy.alloc(n), "alloc");
    HM_TRY(dw.alloc(kModelSize), "alloc");
    HM_TRY(part.alloc((size_t)kGradBlocks * (kModelSize + 1)), "alloc");
    HM_TRY(dg.alloc(kModelSize + 1), "alloc");
    HM_TRY(hipMemcpy(dx.p, x, (size_t)n * kFeatures * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemcpy(dy.p, y, (size_t)n * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemcpy(dw.p, params, kModelSize * 4, hipMemcpyHostToDevice), "copy");
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(kGradBlocks), dim3(256), 0, 0, dx.p, dy.p, nullptr, n,
                       dw.p, pos_weight, part.p);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kModelSize + 256) / 256), dim3(256), 0, 0, part.p,
                       kGradBlocks, n, dg.p);
    HM_TRY(hipGetLastError(), "grad launch");
    std::vector<float> g(kModelSize + 1);
    HM_TRY(hipMemcpy(g.data(), dg.p, (kModelSize + 1) * 4, hipMemcpyDeviceToHost), "copy");
    memcpy(grad, g.data(), kModelSize * 4);
    if (loss) *loss = g[kModelSize];
    return HM_OK;
}

int hm_mlp_forward(const float *x, int64_t n, const float *params, float *logits) {
    if (!x || !params || !logits || n < 0) return invalid("bad arguments");
    if (n == 0) return HM_OK;
    Dev<float> dx, dw, dl;
    HM_TRY(dx.alloc((size_t)n * kFeatures), "alloc");
    HM_TRY(dw.alloc(kModelSize), "alloc");
    HM_TRY(dl.alloc(n), "alloc");
    HM_TRY(hipMemcpy(dx.p, x, (size_t)n * kFeatures * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemcpy(dw.p, params, kModelSize * 4, hipMemcpyHostToDevice), "copy");
    const int blocks = (int)std::min<int64_t>(4096, (n + 3) / 4);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, 0, dx.p, n, dw.p, dl.p);
    HM_TRY(hipGetLastError(), "forward launch");
    HM_TRY(hipMemcpy(logits, dl.p, (size_t)n * 4, hipMemcpyDeviceToHost), "copy");
    return HM_OK;
}

int hm_mlp_train(const float *x, const float *y, int64_t n, float *params, int32_t epochs,
                 int32_t batch, float lr, float pos_weight, uint64_t seed, float *epoch_loss) {
    if (!x || !y || !params || n < 1 || epochs < 1 || batch < 1) return invalid("bad arguments");
    Dev<float> dx, dy, dw, part, dg, dm, dv;
    Dev<int32_t> dperm;
    HM_TRY(dx.alloc((size_t)n * kFeatures), "alloc");
    HM_TRY(dy.alloc(n), "alloc");
    HM_TRY(dw.alloc(kModelSize), "alloc");
    HM_TRY(part.alloc((size_t)kGradBlocks * (kModelSize + 1)), "alloc");
    HM_TRY(dg.alloc(kModelSize + 1), "alloc");
    HM_TRY(dm.alloc(kModelSize), "alloc");
    HM_TRY(dv.alloc(kModelSize), "alloc");
    HM_TRY(dperm.alloc(n), "alloc");
    HM_TRY(hipMemcpy(dx.p, x, (size_t)n * kFeatures * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemcpy(dy.p, y, (size_t)n * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemcpy(dw.p, params, kModelSize * 4, hipMemcpyHostToDevice), "copy");
    HM_TRY(hipMemset(dm.p, 0, kModelSize * 4), "memset");
    HM_TRY(hipMemset(dv.p, 0, kModelSize * 4), "memset");
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    int step = 0;
    std::vector<float> g(kModelSize + 1);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        HM_TRY(hipMemcpy(dperm.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice), "copy");
        double eloss = 0.0;
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int64_t bn = std::min<int64_t>(batch, n - b0);
            const int blocks = (int)std::min<int64_t>(kGradBlocks, (bn + 3) / 4);
            hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, 0, dx.p, dy.p,
                               dperm.p + b0, bn, dw.p, pos_weight, part.p);
            hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kModelSize + 256) / 256), dim3(256), 0, 0,
                               part.p, blocks, bn, dg.p);
            step += 1;
            hipLaunchKernelGGL(adam_kernel, dim3((kModelSize - kModelW1 + 255) / 256), dim3(256), 0,
                               0, dw.p, dg.p, dm.p, dv.p, lr, 0.9f, 0.999f, 1e-8f, step);
            HM_TRY(hipGetLastError(), "train launch");
            if (epoch_loss) {
                float l;
                HM_TRY(hipMemcpy(&l, dg.p + kModelSize, 4, hipMemcpyDeviceToHost), "copy");
                eloss += (double)l * (double)bn;
            }
        }
        if (epoch_loss) epoch_loss[e] = (float)(eloss / (double)n);
    }
    HM_TRY(hipMemcpy(params, dw.p, kModelSize * 4, hipMemcpyDeviceToHost), "copy");
    return HM_OK;
}

int hm_set_stop_model(hm_index *h, const float *params, float threshold, int32_t min_steps) {
    if (!h) return invalid("null argument");
    if (!params) {
        h->stop_on = false;
        return HM_OK;
    }
    if (!(threshold > 0.f && threshold < 1.f)) return invalid("threshold must be in (0, 1)");
    // bfloat16 weights (the search kernel keeps them packed in pairs and normalises the
    // features with the float32 means and inverse standard deviations).
    std::vector<float> w(params, params + kModelSize);
    for (int i = kModelW1; i < kModelSize; ++i) w[i] = bf16_host(w[i]);
    HM_TRY(hipSetDevice(h->device), "hipSetDevice");
    HM_TRY(h->stop_w.ensure(kModelSize), "alloc");
    HM_TRY(hipMemcpy(h->stop_w.p, w.data(), kModelSize * 4, hipMemcpyHostToDevice), "copy model");
    h->stop_logit = logf(threshold / (1.f - threshold));
    h->stop_tmin = std::max(0, (int)min_steps);
    h->stop_on = true;
    return HM_OK;
}

int hm_set_stop_cap(hm_index *h, int32_t max_steps) {
    if (!h) return invalid("null index");
    if (max_steps < 0) return invalid("max_steps must be >= 0");
    h->stop_tmax = max_steps;
    return HM_OK;
}

}  // extern "C"
