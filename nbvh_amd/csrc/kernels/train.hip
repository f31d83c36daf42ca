// train.hip -- one training step of the network on encoded segments and their targets: the
// replacement of the reference's trainer->training_step(stream, inputs[9 x n], targets[8 x n])
// (src/cuda/neural/neural_bvh_renderer.cu:170-173) for its default configuration. FP32 throughout.
//
//   train_forward   one thread per sample: hash-grid features (the query's FP32 grid code) and
//                   the MLP forward pass; features, hidden activations and outputs are stored
//                   per sample (stride kActStride floats) for the backward pass;
//   train_backward  one thread per sample: the composite loss of the reference's defaults
//                   (pt_composite_loss.cuh:19-123, neural_bvh_module.h:250-266), its gradient
//                   through the Sigmoid outputs and the ReLU layers, stored per layer, and the
//                   feature gradients scattered into the table gradient with atomics;
//   train_wgrad     the matrix gradients dW[o][i] = sum over samples of delta[o] * act[i], as
//                   16 x 16 tiles over chunks of samples, added with atomics;
//   train_adam      Adam (Kingma and Ba 2015, Algorithm 1) with L2 regularisation added to the
//                   gradient, on the tables and the matrices; parameters whose gradient is
//                   exactly zero (table entries no sample touched) keep their value and moments.
// Gradients are sums over the batch of the per-element gradients divided by n * 8, as the
// reference's loss kernel writes them; the summation order of the atomics is not fixed.
#include <hip/hip_runtime.h>

#include "../device/common.h"
#include "launch.h"

namespace nbvh {
namespace dev {

constexpr int kK0 = 96, kH = 64, kO = 8, kLayers = 5;
// Per-sample activation record: a0 (features) | h1..h4 | y | d1..d4 | d5.
constexpr int kA0 = 0, kHid = kA0 + kK0, kY = kHid + 4 * kH, kD = kY + kO, kD5 = kD + 4 * kH;
constexpr int kActStride = kD5 + kO;  // 624 floats

__device__ __forceinline__ int w_offset(int layer)
{
    return layer == 0 ? 0 : kH * kK0 + (layer - 1) * kH * kH;
}

template <int P, int L, int F>
__global__ __launch_bounds__(64) void train_forward_kernel(TrainArgs t, const float* __restrict__ in, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float pts[3 * P];
    bool ok = i < n;
#pragma unroll
    for (int c = 0; c < 3 * P; c++) pts[c] = ok ? in[(size_t)i * (3 * P) + c] : 0.5f;
    float x[kK0];
    grid_features<P, L, F, float>(t.net, t.tables, pts, x);  // wave-uniform, every lane takes part
    if (!ok) return;
    float* act = t.act + (size_t)i * kActStride;
#pragma unroll
    for (int k = 0; k < kK0; k++) act[kA0 + k] = x[k];
    // Hidden layers: read the previous activation back from this sample's record.
    const float* prev = act + kA0;
    int fan_in = kK0;
    for (int layer = 0; layer < 4; layer++) {
        const float* W = t.mlp + w_offset(layer);
        float* out = act + kHid + layer * kH;
        for (int o = 0; o < kH; o++) {
            float s = 0.f;
            for (int k = 0; k < fan_in; k++) s = fmaf(W[o * fan_in + k], prev[k], s);
            out[o] = fmaxf(s, 0.f);
        }
        prev = out;
        fan_in = kH;
    }
    const float* W = t.mlp + w_offset(4);
    for (int o = 0; o < kO; o++) {
        float s = 0.f;
        for (int k = 0; k < kH; k++) s = fmaf(W[o * kH + k], prev[k], s);
        act[kY + o] = sigmoidf_cw(s);
    }
}

// Loss and gradient of one output element (the reference's defaults; see oracle/train.py).
__device__ __forceinline__ void element_loss(int j, float y, float target, bool keep, float& loss, float& grad)
{
    loss = 0.f;
    grad = 0.f;
    if (j == 0) {  // BinaryCrossEntropy, prediction clamped to [1e-5, 1 - 1e-5]
        const float p = fminf(fmaxf(y, 1e-5f), 1.f - 1e-5f);
        loss = -(target * logf(p) + (1.f - target) * logf(1.f - p));
        grad = -(target - p) / (p - p * p);
    } else if (!keep) {
        return;  // visibility masking: target visibility < 0.5
    } else if (j == 1) {  // depth: L1
        const float d = y - target;
        loss = fabsf(d);
        grad = copysignf(1.f, d);
    } else if (j < 5) {  // normal: L1 on the linearised output ((y*2-1)/0.1)*0.5+0.5, gradient / 0.1
        const float lin = ((y * 2.f - 1.f) / 0.1f) * 0.5f + 0.5f;
        const float d   = lin - target;
        loss = fabsf(d);
        grad = copysignf(1.f, d) * (1.f / 0.1f);
    } else {  // albedo: relative L2 with the denominator held constant
        const float d   = y - target;
        const float den = y * y + 1e-2f;
        loss = d * d / den;
        grad = 2.f * d / den;
    }
}

template <int P, int L, int F>
__global__ __launch_bounds__(64) void train_backward_kernel(TrainArgs t, const float* __restrict__ in,
                                                            const float* __restrict__ targets, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* act = t.act + (size_t)i * kActStride;
    const float inv_n = 1.f / ((float)n * (float)kO);
    const bool keep = targets[(size_t)i * kO] >= 0.5f;
    float dz[kO], lsum = 0.f;
#pragma unroll
    for (int j = 0; j < kO; j++) {
        const float y = act[kY + j];
        float l, g;
        element_loss(j, y, targets[(size_t)i * kO + j], keep, l, g);
        lsum += l * inv_n;
        dz[j] = (g * inv_n) * (y * (1.f - y));  // through the Sigmoid output activation
        act[kD5 + j] = dz[j];
    }
    atomicAdd(t.loss, lsum);
    // delta of hidden layer 4 from the output matrix, then down the ReLU layers.
    float d[kH];
    {
        const float* W = t.mlp + w_offset(4);
        const float* h = act + kHid + 3 * kH;
        for (int k = 0; k < kH; k++) {
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < kO; o++) s = fmaf(W[o * kH + k], dz[o], s);
            d[k] = h[k] > 0.f ? s : 0.f;
        }
    }
    for (int layer = 3; layer >= 1; layer--) {
        float* dst = act + kD + layer * kH;
        for (int k = 0; k < kH; k++) dst[k] = d[k];
        const float* W = t.mlp + w_offset(layer);
        const float* h = act + kHid + (layer - 1) * kH;
        float nd[kH];
        for (int k = 0; k < kH; k++) {
            float s = 0.f;
            for (int o = 0; o < kH; o++) s = fmaf(W[o * kH + k], d[o], s);
            nd[k] = h[k] > 0.f ? s : 0.f;
        }
        for (int k = 0; k < kH; k++) d[k] = nd[k];
    }
    for (int k = 0; k < kH; k++) act[kD + k] = d[k];
    // Feature gradients and their scatter into the tables' gradient.
    const float* W0 = t.mlp;
    float pts[3 * P];
#pragma unroll
    for (int c = 0; c < 3 * P; c++) pts[c] = in[(size_t)i * (3 * P) + c];
    const bool exact = !in_unit_cube<3 * P>(pts);
    for (int p = 0; p < P; p++) {
        for (int l = 0; l < L; l++) {
            float dx[F];
#pragma unroll
            for (int f = 0; f < F; f++) {
                const int k = p * L * F + l * F + f;
                float s = 0.f;
                for (int o = 0; o < kH; o++) s = fmaf(W0[o * kK0 + k], d[o], s);
                dx[f] = s;
            }
            const GridLevel& lv = t.net.level[l];
            uint32_t g[3], idx[8];
            float fr[3];
            grid_cell(lv, pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], g, fr);
            if (exact)
                grid_corners<true>(lv, g, idx);
            else
                grid_corners<false>(lv, g, idx);
            float* gt = t.grad_tables + (size_t)lv.offset * F;
            for (uint32_t c = 0; c < 8; c++) {
                const float w = corner_weight(c, fr);
#pragma unroll
                for (int f = 0; f < F; f++) atomicAdd(gt + (size_t)idx[c] * F + f, w * dx[f]);
            }
        }
    }
}

// dW[o][i] += sum over samples b in this chunk of delta_b[o] * act_b[i]; 16 x 16 tiles.
__global__ __launch_bounds__(256) void train_wgrad_kernel(TrainArgs t, uint32_t n, int layer, uint32_t chunk)
{
    const int out_w = layer == 4 ? kO : kH, in_w = layer == 0 ? kK0 : kH;
    const int o = blockIdx.x * 16 + (threadIdx.x >> 4), k = blockIdx.y * 16 + (threadIdx.x & 15);
    const int d_off = layer == 4 ? kD5 : kD + layer * kH;
    const int a_off = layer == 0 ? kA0 : kHid + (layer - 1) * kH;
    const uint32_t b0 = blockIdx.z * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
    if (o >= out_w || k >= in_w) return;
    float s = 0.f;
    for (uint32_t b = b0; b < b1; b++) {
        const float* act = t.act + (size_t)b * kActStride;
        s = fmaf(act[d_off + o], act[a_off + k], s);
    }
    atomicAdd(t.grad_mlp + w_offset(layer) + o * in_w + k, s);
}

__global__ __launch_bounds__(256) void train_adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                         float* __restrict__ m, float* __restrict__ v, uint64_t n,
                                                         AdamArgs h)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || g[i] == 0.f) return;  // sparse update: entries no sample touched stay as they are
    const float gi = g[i] + h.l2_reg * w[i];
    const float mi = h.beta1 * m[i] + (1.f - h.beta1) * gi;
    const float vi = h.beta2 * v[i] + (1.f - h.beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - h.lr * (mi * h.inv_bias1) / (sqrtf(vi * h.inv_bias2) + h.epsilon);
}

__global__ __launch_bounds__(256) void train_zero_kernel(float* __restrict__ p, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace dev

size_t train_act_bytes(uint32_t n) { return (size_t)n * dev::kActStride * sizeof(float); }

hipError_t train_zero(float* p, uint64_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::train_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t train_gradients(const TrainArgs& t, const float* in, const float* targets, uint32_t n, hipStream_t stream,
                           bool* supported)
{
    *supported = true;
    if (n == 0) return hipSuccess;
    const NetArgs& a = t.net;
    const dim3 grid((n + 63) / 64), block(64);
#define X(P_, L_, F_)                                                                                       \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_) {                                       \
        hipLaunchKernelGGL((dev::train_forward_kernel<P_, L_, F_>), grid, block, 0, stream, t, in, n);       \
        hipLaunchKernelGGL((dev::train_backward_kernel<P_, L_, F_>), grid, block, 0, stream, t, in, targets, n); \
    } else
    X(3, 16, 2) X(3, 8, 4)
#undef X
    {
        *supported = false;
        return hipSuccess;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t chunk = 1024;
    for (int layer = 0; layer < dev::kLayers; layer++) {
        const int out_w = layer == 4 ? dev::kO : dev::kH, in_w = layer == 0 ? dev::kK0 : dev::kH;
        hipLaunchKernelGGL(dev::train_wgrad_kernel, dim3((out_w + 15) / 16, (in_w + 15) / 16, (n + chunk - 1) / chunk),
                           dim3(256), 0, stream, t, n, layer, chunk);
    }
    return hipGetLastError();
}

hipError_t train_adam(float* w, const float* g, float* m, float* v, uint64_t n, const AdamArgs& h, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::train_adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, w, g, m, v, n, h);
    return hipGetLastError();
}

}  // namespace nbvh
