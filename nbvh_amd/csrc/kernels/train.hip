// train.hip -- one training step of the network on encoded segments and their targets: the
// replacement of the reference's trainer->training_step(stream, inputs[9 x n], targets[8 x n])
// (src/cuda/neural/neural_bvh_renderer.cu:170-173) for its default configuration. FP32 throughout.
//
// Workspace: 624 activation rows of act_stride floats each (row r, sample i at r * stride + i,
// so that a wave's 64 samples touch 256 consecutive bytes of a row), the five matrices transposed,
// and per-chunk partial matrix gradients.
//   train_transpose  W_l [out][in] -> W_l^T [in][out], the backward pass's row order;
//   train_forward    one thread per sample: hash-grid features (the query's FP32 grid code) and
//                    the MLP forward pass; each output neuron is a dot product of a weight row
//                    (uniform across the wave, read through the scalar cache) with the lane's
//                    activations in registers; the layer's outputs go to the activation rows
//                    and to a per-lane LDS column from which the next layer reloads them;
//   train_backward   one thread per sample: the composite loss of the reference's defaults
//                    (pt_composite_loss.cuh:19-123, neural_bvh_module.h:250-266), its gradient
//                    through the Sigmoid outputs and the ReLU layers (rows of W^T), stored per
//                    layer, and the feature gradients scattered into the table gradient with
//                    atomics;
//   train_wgrad      the matrix gradients dW[o][i] = sum over samples of delta[o] * act[i]: each
//                    workgroup reduces one chunk of samples for one matrix through LDS tiles of 32
//                    samples and writes its partial; train_wgrad_reduce adds the partials in
//                    chunk order (deterministic);
//   train_adam       Adam (Kingma and Ba 2015, Algorithm 1) with L2 regularisation added to the
//                    gradient, on the tables and the matrices; parameters whose gradient is
//                    exactly zero (table entries no sample touched) keep their value and moments.
// Gradients are sums over the batch of the per-element gradients divided by n * 8, as the
// reference's loss kernel writes them; the table gradients and the loss are summed with atomics,
// so their summation order is not fixed.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../device/common.h"
#include "launch.h"

namespace nbvh {
namespace dev {

constexpr int kK0 = 96, kH = 64, kO = 8, kLayers = 5;
// Activation rows: a0 (features) | h1..h4 | y | d1..d4 (deltas of h1..h4) | d5 (delta of y).
constexpr int kA0 = 0, kHid = kA0 + kK0, kY = kHid + 4 * kH, kD = kY + kO, kD5 = kD + 4 * kH;
constexpr int kRows   = kD5 + kO;                           // 624 rows
constexpr int kMlp    = kH * kK0 + 3 * kH * kH + kO * kH;  // 18,944 matrix entries
constexpr int kChunks = 256;                                // most matrix-gradient partials

__host__ __device__ constexpr int w_offset(int layer)
{
    return layer == 0 ? 0 : kH * kK0 + (layer - 1) * kH * kH;
}

__device__ __forceinline__ float& row(float* act, size_t stride, int r, uint32_t i)
{
    return act[(size_t)r * stride + i];
}

// W_l^T: entry (k, o) of layer l at w_offset(l) + k * out + o.
__global__ __launch_bounds__(256) void train_transpose_kernel(const float* __restrict__ mlp, float* __restrict__ wt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kMlp) return;
    const int layer = j < w_offset(1) ? 0 : j < w_offset(4) ? 1 + (j - w_offset(1)) / (kH * kH) : 4;
    const int out_w = layer == 4 ? kO : kH, in_w = layer == 0 ? kK0 : kH;
    const int local = j - w_offset(layer), o = local / in_w, k = local % in_w;
    wt[w_offset(layer) + k * out_w + o] = mlp[j];
}

// Every lane reads back only the LDS column it wrote itself (col[k * 64 + lane]): an array with
// run-time indices that stays out of scratch memory, with no barrier needed.
template <int P, int L, int F>
__global__ __launch_bounds__(64) void train_forward_kernel(TrainArgs t, const float* __restrict__ mlp,
                                                           const float* __restrict__ in, float* __restrict__ act,
                                                           uint32_t n)
{
    __shared__ float col[kH * 64];
    const uint32_t lane = threadIdx.x, i = blockIdx.x * 64 + lane;
    const size_t S = t.act_stride;
    float pts[3 * P];
    const bool ok = i < n;
#pragma unroll
    for (int c = 0; c < 3 * P; c++) pts[c] = ok ? in[(size_t)i * (3 * P) + c] : 0.5f;
    float x[kK0];
    grid_features<P, L, F, float>(t.net, t.tables, pts, x);  // wave-uniform, every lane takes part
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < kK0; k++) row(act, S, kA0 + k, i) = x[k];
    for (int o = 0; o < kH; o++) {
        const float* w = mlp + o * kK0;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < kK0; k++) s = fmaf(w[k], x[k], s);
        s = fmaxf(s, 0.f);
        col[o * 64 + lane] = s;
        row(act, S, kHid + o, i) = s;
    }
    float h[kH];
    for (int layer = 1; layer < 4; layer++) {
#pragma unroll
        for (int k = 0; k < kH; k++) h[k] = col[k * 64 + lane];
        const float* W = mlp + w_offset(layer);
        for (int o = 0; o < kH; o++) {
            const float* w = W + o * kH;
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < kH; k++) s = fmaf(w[k], h[k], s);
            s = fmaxf(s, 0.f);
            col[o * 64 + lane] = s;
            row(act, S, kHid + layer * kH + o, i) = s;
        }
    }
#pragma unroll
    for (int k = 0; k < kH; k++) h[k] = col[k * 64 + lane];
    const float* W = mlp + w_offset(4);
#pragma unroll
    for (int o = 0; o < kO; o++) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < kH; k++) s = fmaf(W[o * kH + k], h[k], s);
        row(act, S, kY + o, i) = sigmoidf_cw(s);
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
__global__ __launch_bounds__(64) void train_backward_kernel(TrainArgs t, const float* __restrict__ wt,
                                                            const float* __restrict__ in,
                                                            const float* __restrict__ targets,
                                                            float* __restrict__ act, float* __restrict__ gtab,
                                                            uint32_t n)
{
    __shared__ float col[kH * 64];
    const uint32_t lane = threadIdx.x, i = blockIdx.x * 64 + lane;
    const size_t S = t.act_stride;
    const bool ok = i < n;
    const float inv_n = 1.f / ((float)n * (float)kO);
    float dz[kO], lsum = 0.f;
    if (ok) {
        const bool keep = targets[(size_t)i * kO] >= 0.5f;
#pragma unroll
        for (int j = 0; j < kO; j++) {
            const float y = row(act, S, kY + j, i);
            float l, g;
            element_loss(j, y, targets[(size_t)i * kO + j], keep, l, g);
            lsum += l * inv_n;
            dz[j] = (g * inv_n) * (y * (1.f - y));  // through the Sigmoid output activation
            row(act, S, kD5 + j, i) = dz[j];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) atomicAdd(t.loss, lsum);
    if (!ok) return;
    // delta of h4 from the output matrix (rows of W4^T), then down the ReLU layers.
    {
        const float* T = wt + w_offset(4);
        for (int k = 0; k < kH; k++) {
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < kO; o++) s = fmaf(T[k * kO + o], dz[o], s);
            s = row(act, S, kHid + 3 * kH + k, i) > 0.f ? s : 0.f;
            col[k * 64 + lane] = s;
            row(act, S, kD + 3 * kH + k, i) = s;
        }
    }
    float d[kH];
    for (int layer = 3; layer >= 1; layer--) {
#pragma unroll
        for (int k = 0; k < kH; k++) d[k] = col[k * 64 + lane];
        const float* T = wt + w_offset(layer);
        for (int k = 0; k < kH; k++) {
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < kH; o++) s = fmaf(T[k * kH + o], d[o], s);
            s = row(act, S, kHid + (layer - 1) * kH + k, i) > 0.f ? s : 0.f;
            col[k * 64 + lane] = s;
            row(act, S, kD + (layer - 1) * kH + k, i) = s;
        }
    }
#pragma unroll
    for (int k = 0; k < kH; k++) d[k] = col[k * 64 + lane];  // delta of h1
    // Feature gradients (rows of W0^T) and their scatter into the tables' gradient.
    float pts[3 * P];
#pragma unroll
    for (int c = 0; c < 3 * P; c++) pts[c] = in[(size_t)i * (3 * P) + c];
    const bool exact = !in_unit_cube<3 * P>(pts);
#pragma unroll
    for (int p = 0; p < P; p++) {
        for (int l = 0; l < L; l++) {
            float dx[F];
#pragma unroll
            for (int f = 0; f < F; f++) {
                const float* T = wt + (p * L * F + l * F + f) * kH;
                float s = 0.f;
#pragma unroll
                for (int o = 0; o < kH; o++) s = fmaf(T[o], d[o], s);
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
            float* gt = gtab + (size_t)lv.offset * F;
#pragma unroll
            for (uint32_t c = 0; c < 8; c++) {
                const float w = corner_weight(c, fr);
#pragma unroll
                for (int f = 0; f < F; f++) atomicAdd(gt + (size_t)idx[c] * F + f, w * dx[f]);
            }
        }
    }
}

// Partial dW[o][k] = sum over samples b in [b0, b1) of delta_b[o] * act_b[k] for one matrix
// (O outputs, I inputs): tiles of 32 samples of the O delta rows and the I activation rows are
// staged in LDS (sm, (O + I) x 33 floats), each thread accumulates an MO x MK block.
template <int O, int I>
__device__ __forceinline__ void wgrad_chunk(const float* __restrict__ act, size_t S, int d_row, int a_row, uint32_t b0,
                                            uint32_t b1, float* sm, float* __restrict__ out)
{
    constexpr int TO = O < 16 ? O : 16, TK = 256 / TO, MO = O / TO, MK = I / TK;
    static_assert(TO * MO == O && TK * MK == I, "tile shape");
    const int tid = threadIdx.x, og = tid / TK, kg = tid % TK;
    float acc[MO][MK];
#pragma unroll
    for (int m = 0; m < MO; m++)
#pragma unroll
        for (int q = 0; q < MK; q++) acc[m][q] = 0.f;
    for (uint32_t c0 = b0; c0 < b1; c0 += 32) {
        __syncthreads();
        for (int e = tid; e < (O + I) * 32; e += 256) {
            const int r = e >> 5, j = e & 31;
            const uint32_t b = c0 + j;
            const int src = r < O ? d_row + r : a_row + (r - O);
            sm[r * 33 + j] = b < b1 ? act[(size_t)src * S + b] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int j = 0; j < 32; j++) {
            float dv[MO], av[MK];
#pragma unroll
            for (int m = 0; m < MO; m++) dv[m] = sm[(og * MO + m) * 33 + j];
#pragma unroll
            for (int q = 0; q < MK; q++) av[q] = sm[(O + kg * MK + q) * 33 + j];
#pragma unroll
            for (int m = 0; m < MO; m++)
#pragma unroll
                for (int q = 0; q < MK; q++) acc[m][q] = fmaf(dv[m], av[q], acc[m][q]);
        }
    }
#pragma unroll
    for (int m = 0; m < MO; m++)
#pragma unroll
        for (int q = 0; q < MK; q++) out[(og * MO + m) * I + kg * MK + q] = acc[m][q];
}

// Workgroup (chunk c, matrix l): the partial of samples [c * per, min(n, (c + 1) * per)).
__global__ __launch_bounds__(256) void train_wgrad_kernel(const float* __restrict__ act, uint32_t S, uint32_t n,
                                                          uint32_t per, float* __restrict__ part)
{
    __shared__ float sm[(kH + kK0) * 33];
    const int layer = blockIdx.y;
    const uint32_t b0 = blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
    float* out = part + (size_t)blockIdx.x * kMlp + w_offset(layer);
    if (layer == 0)
        wgrad_chunk<kH, kK0>(act, S, kD, kA0, b0, b1, sm, out);
    else if (layer == 4)
        wgrad_chunk<kO, kH>(act, S, kD5, kHid + 3 * kH, b0, b1, sm, out);
    else
        wgrad_chunk<kH, kH>(act, S, kD + layer * kH, kHid + (layer - 1) * kH, b0, b1, sm, out);
}

// grad[j] = sum of the chunks' partials: eight running sums over chunks c = r mod 8 (eight loads in
// flight per thread), added in a fixed order, so the result does not depend on the schedule.
__global__ __launch_bounds__(64) void train_wgrad_reduce_kernel(const float* __restrict__ part, uint32_t chunks,
                                                                float* __restrict__ grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kMlp) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t c = 0;
    for (; c + 8 <= chunks; c += 8) {
#pragma unroll
        for (int r = 0; r < 8; r++) s[r] += part[(size_t)(c + r) * kMlp + j];
    }
    for (; c < chunks; c++) s[0] += part[(size_t)c * kMlp + j];
    grad[j] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
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

uint32_t train_act_stride(uint32_t max_batch) { return (max_batch + 63u) & ~63u; }

size_t train_act_bytes(uint32_t max_batch)
{
    return ((size_t)dev::kRows * train_act_stride(max_batch) + (size_t)dev::kMlp * (1 + dev::kChunks)) * sizeof(float);
}

hipError_t train_zero(float* p, uint64_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::train_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

// The caller guarantees n <= act_stride (the trainer's max batch).
hipError_t train_gradients(const TrainArgs& t, const float* in, const float* targets, uint32_t n, hipStream_t stream,
                           bool* supported)
{
    *supported = true;
    if (n == 0) return hipSuccess;
    if (n > t.act_stride) return hipErrorInvalidValue;
    const NetArgs& a   = t.net;
    const size_t   S   = t.act_stride;
    float* const   act = t.act;
    float* const   wt  = act + (size_t)dev::kRows * S;
    float* const   part = wt + dev::kMlp;
    const dim3 grid((n + 63) / 64), block(64), grid_mlp((dev::kMlp + 255) / 256);
    bool shape = false;
#define X(P_, L_, F_)                                                                                                \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_) {                                                \
        shape = true;                                                                                                \
        hipLaunchKernelGGL(dev::train_transpose_kernel, grid_mlp, dim3(256), 0, stream, t.mlp, wt);                  \
        hipLaunchKernelGGL((dev::train_forward_kernel<P_, L_, F_>), grid, block, 0, stream, t, t.mlp, in, act, n);    \
        hipLaunchKernelGGL((dev::train_backward_kernel<P_, L_, F_>), grid, block, 0, stream, t, (const float*)wt, in, \
                           targets, act, t.grad_tables, n);                                                          \
    }
    X(3, 16, 2) X(3, 8, 4)
#undef X
    if (!shape) {
        *supported = false;
        return hipSuccess;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // Chunks of a multiple of 32 samples, at most kChunks of them.
    const uint32_t target = std::min<uint32_t>(dev::kChunks, (n + 255) / 256);
    const uint32_t per = ((n + target - 1) / target + 31u) & ~31u, chunks = (n + per - 1) / per;
    hipLaunchKernelGGL(dev::train_wgrad_kernel, dim3(chunks, dev::kLayers), dim3(256), 0, stream, (const float*)act,
                       (uint32_t)S, n, per, part);
    hipLaunchKernelGGL(dev::train_wgrad_reduce_kernel, dim3((dev::kMlp + 63) / 64), dim3(64), 0, stream,
                       (const float*)part, chunks, t.grad_mlp);
    return hipGetLastError();
}

hipError_t train_adam(float* w, const float* g, float* m, float* v, uint64_t n, const AdamArgs& h, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::train_adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, w, g, m, v, n, h);
    return hipGetLastError();
}

}  // namespace nbvh
