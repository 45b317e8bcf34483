// mlp.hip -- the learned step-size proposer: a 12 -> 64 -> 64 -> 1 MLP with
// tanh hidden layers, BF16 weights (FP32 master copies for Adam) and FP32
// accumulation.  Forward, backward (per-block gradient partials reduced in a
// fixed order, so gradients are deterministic) and Adam run on the GPU.
//
// Parameter vector layout (FP32 master / BF16 working copy), kNParam floats:
//   W1[12][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/red_abi.h"
#include "device_scope.hpp"
#include "mlp.hpp"

namespace rdamd {

constexpr int kIn = RD_MLP_IN, kH = RD_MLP_HIDDEN;
constexpr int kOffW1 = 0, kOffB1 = kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;
constexpr int kNParam = kOffB3 + 1;
static_assert(kNParam == RD_MLP_NPARAM, "parameter count");
constexpr int kMlpBlock = 64;

__device__ inline float bf16_to_f32(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}
__host__ __device__ inline uint16_t f32_to_bf16(float f)
{
    // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // NaN
    const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
    return (uint16_t)((u + r) >> 16);
}

// Standardisation and target.  The network's output y' is de-normalised to
// y' t_std + t_mean + x[base] (base < 0: no offset), so a network trained on
// targets t - x[base] still returns log h.  quantile 0 trains the mean
// (loss 0.5 e^2); quantile q in (0, 1) trains the q-quantile (pinball loss
// max(q u, (q - 1) u), u = t' - y').
struct Norm {
    float mean[kIn], inv_std[kIn];
    float t_mean, t_std;
    int base;
    float quantile;
};

// Weights of one block, expanded to FP32 in LDS.
struct SharedW {
    float W1[kIn * kH], b1[kH], W2[kH * kH], b2[kH], w3[kH], b3;
};

__device__ void load_weights(SharedW& s, const uint16_t* __restrict__ wb)
{
    float* dst = &s.W1[0];
    for (int k = threadIdx.x; k < kNParam; k += blockDim.x) dst[k] = bf16_to_f32(wb[k]);
}

// y = w3 . tanh(W2^T tanh(W1^T x + b1) + b2) + b3 for one normalised input.
__device__ float forward_one(const SharedW& s, const float* x, float* h1, float* h2)
{
    for (int j = 0; j < kH; ++j) {
        float a = s.b1[j];
        for (int i = 0; i < kIn; ++i) a = fmaf(x[i], s.W1[i * kH + j], a);
        h1[j] = tanhf(a);
    }
    for (int j = 0; j < kH; ++j) {
        float a = s.b2[j];
        for (int i = 0; i < kH; ++i) a = fmaf(h1[i], s.W2[i * kH + j], a);
        h2[j] = tanhf(a);
    }
    float y = s.b3;
    for (int j = 0; j < kH; ++j) y = fmaf(h2[j], s.w3[j], y);
    return y;
}

// Inference: raw features -> log h (de-normalised).  One wave per sample,
// lane j = hidden unit j: layer 1 and layer 2 are 12- and 64-term dot
// products per lane over inputs held in LDS; the output sums the 64 lanes'
// w3_j h2_j with a fixed butterfly, so results are reproducible.
__global__ __launch_bounds__(kH) void k_mlp_forward(const uint16_t* __restrict__ wb, Norm nm,
                                                    const float* __restrict__ xraw, uint32_t n,
                                                    float* __restrict__ y)
{
    static_assert(kH == 64, "one lane per hidden unit of a 64-wide wave");
    __shared__ float xs[kIn];
    __shared__ float h1[kH];
    const uint32_t q = blockIdx.x;
    const uint32_t j = threadIdx.x;
    if (q >= n) return;
    if (j < (uint32_t)kIn) xs[j] = (xraw[(size_t)q * kIn + j] - nm.mean[j]) * nm.inv_std[j];
    __syncthreads();
    float a = bf16_to_f32(wb[kOffB1 + j]);
#pragma unroll
    for (int i = 0; i < kIn; ++i) a = fmaf(xs[i], bf16_to_f32(wb[kOffW1 + i * kH + j]), a);
    h1[j] = tanhf(a);
    __syncthreads();
    float b = bf16_to_f32(wb[kOffB2 + j]);
#pragma unroll 16
    for (int k = 0; k < kH; ++k) b = fmaf(h1[k], bf16_to_f32(wb[kOffW2 + k * kH + j]), b);
    float o = tanhf(b) * bf16_to_f32(wb[kOffW3 + j]);
    for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
    if (j == 0) {
        const float off = nm.base >= 0 ? xraw[(size_t)q * kIn + nm.base] : 0.0f;
        y[q] = (o + bf16_to_f32(wb[kOffB3])) * nm.t_std + nm.t_mean + off;
    }
}

// Loss in normalised target units -- 0.5 mean e^2, or the mean pinball loss
// when nm.quantile > 0 (e = y' - t') -- and its gradient.
// Each block (64 samples) writes its partial gradient (and loss) to
// part[block][*]; k_mlp_reduce sums the partials in block order.  Outer
// products over the block's samples go through two LDS buffers in phases
// (W2/b2, then W1/b1, then w3/b3) so the block stays within 64 KB of LDS.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_grad(const uint16_t* __restrict__ wb, Norm nm,
                                                        const float* __restrict__ xraw, const float* __restrict__ tr,
                                                        const uint32_t* __restrict__ idx, uint32_t n, float scale,
                                                        float* __restrict__ part)
{
    __shared__ SharedW s;
    __shared__ float bufA[kMlpBlock][kH + 1];
    __shared__ float bufB[kMlpBlock][kH + 1];
    __shared__ float sx[kMlpBlock][kIn + 1];
    __shared__ float sdy[kMlpBlock];
    __shared__ float sloss[kMlpBlock];
    load_weights(s, wb);
    __syncthreads();
    const uint32_t t = threadIdx.x;
    const uint32_t q = blockIdx.x * kMlpBlock + t;
    const bool valid = q < n;
    float x[kIn], h1[kH], h2[kH];
    float dy = 0.0f, l = 0.0f;
    if (valid) {
        const uint32_t src = idx ? idx[q] : q;
        for (int i = 0; i < kIn; ++i) x[i] = (xraw[(size_t)src * kIn + i] - nm.mean[i]) * nm.inv_std[i];
        const float y = forward_one(s, x, h1, h2);
        const float tb = tr[src] - (nm.base >= 0 ? xraw[(size_t)src * kIn + nm.base] : 0.0f);
        const float e = y - (tb - nm.t_mean) / nm.t_std;
        if (nm.quantile > 0.0f) {
            const float w = e > 0.0f ? 1.0f - nm.quantile : -nm.quantile;
            dy = w * scale;
            l = w * e * scale;
        } else {
            dy = e * scale;
            l = 0.5f * e * e * scale;
        }
    } else {
        for (int i = 0; i < kIn; ++i) x[i] = 0.0f;
        for (int j = 0; j < kH; ++j) h1[j] = h2[j] = 0.0f;
    }
    float* out = part + (size_t)blockIdx.x * (kNParam + 1);
    // phase 1: bufA = h1, bufB = d2 = dy w3 (1 - h2^2); W2[i][j] = sum h1_i d2_j, b2 = sum d2
    for (int j = 0; j < kH; ++j) {
        bufA[t][j] = h1[j];
        bufB[t][j] = dy * s.w3[j] * (1.0f - h2[j] * h2[j]);
    }
    for (int i = 0; i < kIn; ++i) sx[t][i] = x[i];
    sdy[t] = dy;
    sloss[t] = l;
    __syncthreads();
    for (int k = t; k < kH * kH; k += kMlpBlock) {
        const int i = k / kH, j = k % kH;
        float a = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) a = fmaf(bufA[r][i], bufB[r][j], a);
        out[kOffW2 + k] = a;
    }
    for (int j = t; j < kH; j += kMlpBlock) {
        float a = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) a += bufB[r][j];
        out[kOffB2 + j] = a;
    }
    // d1 = (W2 d2) (1 - h1^2), own row of bufB
    float d1[kH];
    for (int i = 0; i < kH; ++i) {
        float a = 0.0f;
        for (int j = 0; j < kH; ++j) a = fmaf(s.W2[i * kH + j], bufB[t][j], a);
        d1[i] = a * (1.0f - h1[i] * h1[i]);
    }
    __syncthreads();
    // phase 2: bufA = d1; W1[i][j] = sum x_i d1_j, b1 = sum d1
    for (int j = 0; j < kH; ++j) bufA[t][j] = d1[j];
    __syncthreads();
    for (int k = t; k < kIn * kH; k += kMlpBlock) {
        const int i = k / kH, j = k % kH;
        float a = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) a = fmaf(sx[r][i], bufA[r][j], a);
        out[kOffW1 + k] = a;
    }
    for (int j = t; j < kH; j += kMlpBlock) {
        float a = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) a += bufA[r][j];
        out[kOffB1 + j] = a;
    }
    __syncthreads();
    // phase 3: bufA = h2; w3[j] = sum dy h2_j, b3 = sum dy
    for (int j = 0; j < kH; ++j) bufA[t][j] = h2[j];
    __syncthreads();
    for (int j = t; j < kH; j += kMlpBlock) {
        float a = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) a = fmaf(sdy[r], bufA[r][j], a);
        out[kOffW3 + j] = a;
    }
    if (t == 0) {
        float a = 0.0f, b = 0.0f;
        for (int r = 0; r < kMlpBlock; ++r) {
            a += sdy[r];
            b += sloss[r];
        }
        out[kOffB3] = a;
        out[kNParam] = b;
    }
}

// Sums per-block partials in block order: out[k] = sum_b part[b][k].
__global__ void k_mlp_reduce(const float* __restrict__ part, uint32_t nblk, float* __restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > (uint32_t)kNParam) return;
    float a = 0.0f;
    for (uint32_t b = 0; b < nblk; ++b) a += part[(size_t)b * (kNParam + 1) + k];
    out[k] = a;
}

// Adam on the FP32 master weights, then the BF16 working copy.
__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, uint16_t* __restrict__ wb, float lr, float b1, float b2,
                           float c1, float c2)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint32_t)kNParam) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
    w[k] = wk;
    wb[k] = f32_to_bf16(wk);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb, uint32_t n)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) wb[k] = f32_to_bf16(w[k]);
}

}  // namespace rdamd

using namespace rdamd;

struct rd_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *part = nullptr;
    uint16_t* wb = nullptr;
    size_t part_blocks = 0;
    float *dx = nullptr, *dt = nullptr, *dy = nullptr;
    uint32_t* didx = nullptr;
    size_t cap = 0;
    Norm nm{};
    uint64_t adam_t = 0;
    std::mt19937 rng;
};

namespace {
thread_local std::string g_mlp_err;
int mfail(int code, const std::string& s)
{
    g_mlp_err = s;
    rdamd::set_last_error(s);
    return code;
}
#define MLP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return mfail(RD_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure_cap(rd_mlp* p, size_t n)
{
    if (n <= p->cap) return RD_OK;
    if (p->dx) (void)hipFree(p->dx);
    if (p->dt) (void)hipFree(p->dt);
    if (p->dy) (void)hipFree(p->dy);
    if (p->didx) (void)hipFree(p->didx);
    MLP_TRY(hipMalloc((void**)&p->dx, n * kIn * sizeof(float)));
    MLP_TRY(hipMalloc((void**)&p->dt, n * sizeof(float)));
    MLP_TRY(hipMalloc((void**)&p->dy, n * sizeof(float)));
    MLP_TRY(hipMalloc((void**)&p->didx, n * sizeof(uint32_t)));
    p->cap = n;
    return RD_OK;
}
int ensure_part(rd_mlp* p, size_t nblk)
{
    if (nblk <= p->part_blocks) return RD_OK;
    if (p->part) (void)hipFree(p->part);
    MLP_TRY(hipMalloc((void**)&p->part, nblk * (kNParam + 1) * sizeof(float)));
    p->part_blocks = nblk;
    return RD_OK;
}
}  // namespace

namespace rdamd {
const uint16_t* mlp_weights(const rd_mlp* p) { return p->wb; }
int mlp_device(const rd_mlp* p) { return p->device; }
int mlp_forward_dev(const rd_mlp* p, hipStream_t st, const float* x, uint32_t n, float* y)
{
    if (n == 0) return RD_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3(n), dim3(kH), 0, st, p->wb, p->nm, x, n, y);
    return hipGetLastError() == hipSuccess ? RD_OK : RD_EDEVICE;
}
}  // namespace rdamd

extern "C" {

int rd_mlp_create(int device, uint32_t seed, rd_mlp** out)
{
    DeviceScope dev_scope(device);
    if (!out) return mfail(RD_EINVAL, "NULL argument");
    rd_mlp* p = new rd_mlp();
    p->device = device;
    p->rng.seed(seed);
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete p;
        return mfail(RD_EDEVICE, std::string("rd_mlp_create: ") + hipGetErrorString(e));
    }
    e = hipMalloc((void**)&p->w, kNParam * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&p->m, kNParam * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&p->v, kNParam * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&p->g, (kNParam + 1) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&p->wb, kNParam * sizeof(uint16_t));
    if (e != hipSuccess) {
        rd_mlp_destroy(p);
        return mfail(RD_ENOMEM, std::string("rd_mlp_create: hipMalloc: ") + hipGetErrorString(e));
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> w(kNParam, 0.0f);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const float lim = std::sqrt(6.0f / (fan_in + fan_out));
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int k = 0; k < count; ++k) w[off + k] = u(p->rng);
    };
    fill(kOffW1, kIn, kH, kIn * kH);
    fill(kOffW2, kH, kH, kH * kH);
    fill(kOffW3, kH, 1, kH);
    for (int i = 0; i < kIn; ++i) { p->nm.mean[i] = 0.0f; p->nm.inv_std[i] = 1.0f; }
    p->nm.t_mean = 0.0f;
    p->nm.t_std = 1.0f;
    p->nm.base = -1;
    p->nm.quantile = 0.0f;
    if (int rc = rd_mlp_set_params(p, w.data())) {
        rd_mlp_destroy(p);
        return rc;
    }
    *out = p;
    return RD_OK;
}

int rd_mlp_destroy(rd_mlp* p)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p) return RD_OK;
    (void)hipStreamSynchronize(p->stream);
    for (void* q : {(void*)p->w, (void*)p->m, (void*)p->v, (void*)p->g, (void*)p->wb, (void*)p->part, (void*)p->dx,
                    (void*)p->dt, (void*)p->dy, (void*)p->didx})
        if (q) (void)hipFree(q);
    (void)hipStreamDestroy(p->stream);
    delete p;
    return RD_OK;
}

int rd_mlp_set_norm(rd_mlp* p, const float mean[RD_MLP_IN], const float stdev[RD_MLP_IN], float t_mean, float t_std)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p || !mean || !stdev || !(t_std > 0.0f)) return mfail(RD_EINVAL, "bad argument");
    for (int i = 0; i < kIn; ++i) {
        if (!(stdev[i] > 0.0f)) return mfail(RD_EINVAL, "feature std must be > 0");
        p->nm.mean[i] = mean[i];
        p->nm.inv_std[i] = 1.0f / stdev[i];
    }
    p->nm.t_mean = t_mean;
    p->nm.t_std = t_std;
    return RD_OK;
}

int rd_mlp_set_target(rd_mlp* p, int base_feature, float quantile)
{
    if (!p) return mfail(RD_EINVAL, "NULL model");
    if (base_feature < -1 || base_feature >= kIn) return mfail(RD_EINVAL, "base feature must be -1 or 0..11");
    if (!(quantile == 0.0f || (quantile > 0.0f && quantile < 1.0f)))
        return mfail(RD_EINVAL, "quantile must be 0 (mean) or in (0, 1)");
    p->nm.base = base_feature;
    p->nm.quantile = quantile;
    return RD_OK;
}

int rd_mlp_set_params(rd_mlp* p, const float* params)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p || !params) return mfail(RD_EINVAL, "bad argument");
    MLP_TRY(hipMemcpyAsync(p->w, params, kNParam * sizeof(float), hipMemcpyHostToDevice, p->stream));
    MLP_TRY(hipMemsetAsync(p->m, 0, kNParam * sizeof(float), p->stream));
    MLP_TRY(hipMemsetAsync(p->v, 0, kNParam * sizeof(float), p->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((kNParam + 255) / 256), dim3(256), 0, p->stream, p->w, p->wb,
                       (uint32_t)kNParam);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(p->stream));
    p->adam_t = 0;
    return RD_OK;
}

int rd_mlp_get_params(rd_mlp* p, float* params, uint16_t* params_bf16)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p) return mfail(RD_EINVAL, "bad argument");
    if (params) MLP_TRY(hipMemcpyAsync(params, p->w, kNParam * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    if (params_bf16)
        MLP_TRY(hipMemcpyAsync(params_bf16, p->wb, kNParam * sizeof(uint16_t), hipMemcpyDeviceToHost, p->stream));
    MLP_TRY(hipStreamSynchronize(p->stream));
    return RD_OK;
}

int rd_mlp_forward(rd_mlp* p, const float* x, uint32_t n, float* y)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p || (!x && n) || (!y && n)) return mfail(RD_EINVAL, "bad argument");
    if (n == 0) return RD_OK;
    if (int rc = ensure_cap(p, n)) return rc;
    MLP_TRY(hipMemcpyAsync(p->dx, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, p->stream));
    if (int rc = mlp_forward_dev(p, p->stream, p->dx, n, p->dy)) return mfail(rc, "k_mlp_forward launch");
    MLP_TRY(hipMemcpyAsync(y, p->dy, n * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    MLP_TRY(hipStreamSynchronize(p->stream));
    return RD_OK;
}

static int grad_dev(rd_mlp* p, const uint32_t* idx, uint32_t n, float scale)
{
    const uint32_t nblk = (n + kMlpBlock - 1) / kMlpBlock;
    if (int rc = ensure_part(p, nblk)) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(kMlpBlock), 0, p->stream, p->wb, p->nm, p->dx, p->dt, idx, n,
                       scale, p->part);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 1 + 255) / 256), dim3(256), 0, p->stream, p->part, nblk, p->g);
    MLP_TRY(hipGetLastError());
    return RD_OK;
}

int rd_mlp_loss_grad(rd_mlp* p, const float* x, const float* t, uint32_t n, float* loss, float* grad)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p || !x || !t || n == 0) return mfail(RD_EINVAL, "bad argument");
    if (int rc = ensure_cap(p, n)) return rc;
    MLP_TRY(hipMemcpyAsync(p->dx, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, p->stream));
    MLP_TRY(hipMemcpyAsync(p->dt, t, n * sizeof(float), hipMemcpyHostToDevice, p->stream));
    if (int rc = grad_dev(p, nullptr, n, 1.0f / n)) return rc;
    std::vector<float> g(kNParam + 1);
    MLP_TRY(hipMemcpyAsync(g.data(), p->g, (kNParam + 1) * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    MLP_TRY(hipStreamSynchronize(p->stream));
    if (grad) std::memcpy(grad, g.data(), kNParam * sizeof(float));
    if (loss) *loss = g[kNParam];
    return RD_OK;
}

int rd_mlp_train(rd_mlp* p, const float* x, const float* t, uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                 float* final_loss)
{
    DeviceScope dev_scope(p ? p->device : -1);
    if (!p || !x || !t || n == 0 || batch == 0) return mfail(RD_EINVAL, "bad argument");
    if (int rc = ensure_cap(p, n)) return rc;
    MLP_TRY(hipMemcpyAsync(p->dx, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, p->stream));
    MLP_TRY(hipMemcpyAsync(p->dt, t, n * sizeof(float), hipMemcpyHostToDevice, p->stream));
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    const float b1 = 0.9f, b2 = 0.999f;
    double last = 0.0;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), p->rng);
        MLP_TRY(hipMemcpyAsync(p->didx, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, p->stream));
        double ep_loss = 0.0;
        std::vector<float> lbuf;
        for (uint32_t s = 0; s < n; s += batch) {
            const uint32_t nb = std::min(batch, n - s);
            if (int rc = grad_dev(p, p->didx + s, nb, 1.0f / nb)) return rc;
            p->adam_t++;
            const float c1 = 1.0f - std::pow(b1, (float)p->adam_t), c2 = 1.0f - std::pow(b2, (float)p->adam_t);
            hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, p->stream, p->w, p->m, p->v,
                               p->g, p->wb, lr, b1, b2, c1, c2);
            MLP_TRY(hipGetLastError());
            if (ep + 1 == epochs) {
                float l;
                MLP_TRY(hipMemcpyAsync(&l, p->g + kNParam, sizeof(float), hipMemcpyDeviceToHost, p->stream));
                MLP_TRY(hipStreamSynchronize(p->stream));
                ep_loss += (double)l * nb;
            }
        }
        MLP_TRY(hipStreamSynchronize(p->stream));
        last = ep_loss / n;
    }
    if (final_loss) *final_loss = (float)last;
    return RD_OK;
}

}  // extern "C"
