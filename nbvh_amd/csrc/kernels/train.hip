// train.hip -- one training step of the network on encoded segments and their targets: the
// replacement of the reference's trainer->training_step(stream, inputs[9 x n], targets[8 x n])
// (src/cuda/neural/neural_bvh_renderer.cu:170-173) for its default configuration. FP32 throughout.
//
// Workspace: 720 activation rows of act_stride floats each (row r, sample i at r * stride + i,
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
//                    layer, down to the feature gradients;
//   train_scatter    the table gradient of levels of fewer than 128 slices of 128 KB, one
//                    workgroup per slice: the slice is summed in LDS from every (sample, point)
//                    whose corners fall in it and stored whole;
//   train_bin_*      the table gradient of the other levels: the corner terms binned by slice
//                    (count, scan, scatter), then one workgroup per slice sums its bin in LDS
//                    and stores the slice whole; no global atomics on the gradient;
//   train_wgrad      the matrix gradients dW[o][i] = sum over samples of delta[o] * act[i]: each
//                    workgroup reduces one chunk of samples for one matrix through LDS tiles of 32
//                    samples and writes its partial; train_wgrad_reduce adds the partials in
//                    chunk order (deterministic);
//   train_adam       tiny-cuda-nn's Adam: per-parameter update counts for the bias correction,
//                    L2 regularisation and dense updates for the matrices, sparse updates
//                    without L2 for the tables (entries no sample touched keep their value).
// Gradients are sums over the batch of the per-element gradients divided by n * 8, as the
// reference's loss kernel writes them; the table gradients (LDS atomics) and the loss are summed
// in an order that is not fixed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../device/common.h"
#include "launch.h"

namespace nbvh {
namespace dev {

constexpr int kK0 = 96, kH = 64, kO = 8, kLayers = 5;
// Activation rows: a0 (features) | h1..h4 | y | d1..d4 (deltas of h1..h4) | d5 (delta of y) |
// dx (feature gradients).
constexpr int kA0 = 0, kHid = kA0 + kK0, kY = kHid + 4 * kH, kD = kY + kO, kD5 = kD + 4 * kH, kDX = kD5 + kO;
constexpr int kRows   = kDX + kK0;                          // 720 rows
#ifndef NBVH_SCATTER_SLICE_FLOATS
#define NBVH_SCATTER_SLICE_FLOATS 32768
#endif
constexpr int kSliceFloats = NBVH_SCATTER_SLICE_FLOATS;     // table-gradient slice in LDS (128 KB)
constexpr uint32_t kScatterThreads = 1024;                  // threads sharing one slice
#ifndef NBVH_SCATTER_UNROLL
#define NBVH_SCATTER_UNROLL 2                               // (sample, point) pairs per thread and iteration
#endif
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
                                                            float* __restrict__ act, uint32_t n)
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
        if (t.sample_loss) t.sample_loss[i] = lsum;  // this sample's element values, in element order
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
    // Feature gradients (rows of W0^T); train_scatter adds them into the tables' gradient.
#pragma unroll 4
    for (int k = 0; k < kK0; k++) {
        const float* T = wt + k * kH;
        float s = 0.f;
#pragma unroll
        for (int o = 0; o < kH; o++) s = fmaf(T[o], d[o], s);
        row(act, S, kDX + k, i) = s;
    }
}

// Corner entries of a hashed level: grid_corners' coherent-prime hash alone.
__device__ __forceinline__ void hash_corners(const GridLevel& L, const uint32_t g[3], uint32_t idx[8])
{
    const uint32_t hy0 = g[1] * 2654435761u, hy1 = (g[1] + 1u) * 2654435761u;
    const uint32_t hz0 = g[2] * 805459861u, hz1 = (g[2] + 1u) * 805459861u;
#pragma unroll
    for (uint32_t c = 0; c < 8; c++)
        idx[c] = ((g[0] + (c & 1u)) ^ ((c & 2u) ? hy1 : hy0) ^ ((c & 4u) ? hz1 : hz0)) & (L.size - 1u);
}

// Work items of the table gradient. Slice j of level l holds the level's entries
// [j * E, min((j + 1) * E, size)), E = kSliceFloats / F; each slice is split into parts[l] parts
// of the batch's (sample, point) pairs; level l's items are [first[l], first[l + 1]), item
// first[l] + j * parts[l] + k being part k of slice j.
struct SliceMap {
    uint32_t first[kMaxLevels + 1];
    uint32_t parts[kMaxLevels];
};

// One workgroup per item: the slice's gradient accumulates in LDS while the workgroup visits its
// part of the batch's (sample, point) pairs, recomputes the cell at the slice's level and adds
// w_c * dx[f] for each corner c whose entry lies in the slice (LDS atomics). A slice of a single
// part is then stored whole, zeros included; the parts of a split slice add their non-zero
// entries to the (zeroed) gradient with global atomics. Every entry of every level belongs to
// exactly one slice, so the table gradient is written completely. The cost is a scan of the
// batch's points per item instead of one global atomic add per (point, level, corner, feature);
// slices of the coarse dense levels, which receive many corners, are split so that no workgroup
// adds many more corners than the others.
template <int P, int L, int F>
__global__ __launch_bounds__(kScatterThreads) void train_scatter_kernel(TrainArgs t, SliceMap map,
                                                                        const float* __restrict__ in,
                                                                        const float* __restrict__ act,
                                                                        float* __restrict__ gtab, uint32_t n)
{
    constexpr uint32_t E = kSliceFloats / F;
    __shared__ float acc[kSliceFloats];
    int l = 0;
    while (l + 1 < L && blockIdx.x >= map.first[l + 1]) l++;
    const GridLevel& lv = t.net.level[l];
    const uint32_t parts = map.parts[l], item = blockIdx.x - map.first[l];
    const uint32_t slice = item / parts, part = item - slice * parts;
    const uint32_t e0 = slice * E;
    const uint32_t len = lv.size - e0 < E ? lv.size - e0 : E;
    for (uint32_t j = threadIdx.x; j < len * F; j += kScatterThreads) acc[j] = 0.f;
    __syncthreads();
    const size_t S = t.act_stride;
    const uint32_t total = n * (uint32_t)P;
    const uint32_t j0 = (uint32_t)((uint64_t)total * part / parts), j1 = (uint32_t)((uint64_t)total * (part + 1) / parts);
    // U (sample, point) pairs per thread and iteration, their point loads issued together.
    constexpr uint32_t B = kScatterThreads;
    constexpr int U = NBVH_SCATTER_UNROLL;
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += U * B) {
        float pts[U][3];
#pragma unroll
        for (int h = 0; h < U; h++) {
            const uint32_t jj = j + h * B < j1 ? j + h * B : j;
            const uint32_t i = jj / P, p = jj - i * P;
#pragma unroll
            for (int c = 0; c < 3; c++) pts[h][c] = in[(size_t)i * (3 * P) + 3 * p + c];
        }
#pragma unroll
        for (int h = 0; h < U; h++) {
            const uint32_t jj = j + h * B;
            if (jj >= j1) break;
            const uint32_t i = jj / P, p = jj - i * P;
            uint32_t g[3], idx[8];
            float fr[3];
            grid_cell(lv, pts[h][0], pts[h][1], pts[h][2], g, fr);
            if (!lv.dense)  // uniform per workgroup: only the level's own index formula
                hash_corners(lv, g, idx);
            else if (in_unit_cube<3>(pts[h]))
                grid_corners<false>(lv, g, idx);
            else
                grid_corners<true>(lv, g, idx);
            bool any = false;
#pragma unroll
            for (int c = 0; c < 8; c++) any = any || (idx[c] - e0 < len);
            if (!any) continue;
            float dx[F];
#pragma unroll
            for (int f = 0; f < F; f++) dx[f] = act[(size_t)(kDX + (p * L + l) * F + f) * S + i];
#pragma unroll
            for (uint32_t c = 0; c < 8; c++) {
                const uint32_t r = idx[c] - e0;
                if (r < len) {
                    const float w = corner_weight(c, fr);
#pragma unroll
                    for (int f = 0; f < F; f++) atomicAdd(&acc[r * F + f], w * dx[f]);
                }
            }
        }
    }
    __syncthreads();
    float* out = gtab + ((size_t)lv.offset + e0) * F;
    if (parts == 1) {
        for (uint32_t j = threadIdx.x; j < len * F; j += B) out[j] = acc[j];
    } else {
        for (uint32_t j = threadIdx.x; j < len * F; j += B)
            if (acc[j] != 0.f) atomicAdd(out + j, acc[j]);
    }
}

// Binned table gradient, for levels of 128 or more 128 KB slices: the level is cut into bins of
// kBinFloats gradient floats; every corner term (entry, w_c * dx[f]) of the level is written to
// its bin, then one workgroup per bin adds the bin's terms in LDS and stores the bin whole. Four
// passes:
//   train_bin_count_kernel    per workgroup a histogram of its pairs' corner bins in LDS, added to
//                             the bin counts (one global atomic per workgroup and bin touched);
//   train_bin_scan_kernel     exclusive scan of the counts: bin b's terms at [offsets[b],
//                             offsets[b + 1]); the cursors start at the offsets;
//   train_bin_scatter_kernel  the same pairs again: each workgroup reserves its range of every
//                             bin (one global atomic per bin touched) and writes its terms;
//   train_bin_accumulate_kernel  one workgroup per bin: LDS sum, the bin stored whole.
// The batch is read three times and each term twice, instead of one global atomic add per
// (pair, corner, feature) or one scan of the batch per slice.
struct BinMap {
    uint32_t count;                    // binned levels
    uint32_t level[kMaxLevels];        // their indices
    uint32_t first[kMaxLevels + 1];    // first bin of binned level k; first[count] = bins
};
#ifndef NBVH_BIN_THREADS
#define NBVH_BIN_THREADS 256
#endif
// 8192 pairs: a workgroup's terms reach each bin in runs long enough to write coalesced. Measured
// per step at 2^21 / 2^22 / 2^23 entries (256 threads): 2048 pairs 1.74 / 2.14 / 2.68 ms, 8192
// pairs 1.61 / 2.02 / 2.56 ms (1024 threads: 1.59 / 2.07 / 2.62).
#ifndef NBVH_BIN_PAIRS
#define NBVH_BIN_PAIRS 8192
#endif
constexpr uint32_t kBinThreads = NBVH_BIN_THREADS;
constexpr uint32_t kBinPairs   = NBVH_BIN_PAIRS;  // (sample, point) pairs per workgroup and level
constexpr uint32_t kBinFloats  = 16384;   // gradient floats per bin (64 KB of LDS: 2 workgroups per CU)
constexpr uint32_t kBinWindow  = 8192;    // bins of one level a count / scatter workgroup handles
constexpr uint32_t kBinUnroll  = 4;       // terms per thread whose loads are issued together

// Corners and fractional position of pair jj = (sample i, point p) at level lv.
template <int P>
__device__ __forceinline__ void pair_corners(const GridLevel& lv, const float* __restrict__ in, uint32_t i, uint32_t p,
                                             uint32_t idx[8], float fr[3])
{
    float pt[3];
#pragma unroll
    for (int c = 0; c < 3; c++) pt[c] = in[(size_t)i * (3 * P) + 3 * p + c];
    uint32_t g[3];
    grid_cell(lv, pt[0], pt[1], pt[2], g, fr);
    if (!lv.dense)
        hash_corners(lv, g, idx);
    else if (in_unit_cube<3>(pt))
        grid_corners<false>(lv, g, idx);
    else
        grid_corners<true>(lv, g, idx);
}

// Workgroup (x, k, w): pairs [x * kBinPairs, ...) of binned level k, the corners in the level's
// bins [w * kBinWindow, (w + 1) * kBinWindow); hist: one counter per bin of the window (dynamic
// LDS).
template <int P, int F>
__global__ __launch_bounds__(kBinThreads) void train_bin_count_kernel(TrainArgs t, BinMap bm,
                                                                      const float* __restrict__ in, uint32_t n,
                                                                      uint32_t* __restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    constexpr uint32_t E = kBinFloats / F;
    const uint32_t k = blockIdx.y, w0 = blockIdx.z * kBinWindow, level_bins = bm.first[k + 1] - bm.first[k];
    if (w0 >= level_bins) return;
    const uint32_t first = bm.first[k] + w0, bins = level_bins - w0 < kBinWindow ? level_bins - w0 : kBinWindow;
    const GridLevel& lv = t.net.level[bm.level[k]];
    for (uint32_t j = threadIdx.x; j < bins; j += kBinThreads) hist[j] = 0u;
    __syncthreads();
    const uint32_t total = n * (uint32_t)P, j0 = blockIdx.x * kBinPairs;
    const uint32_t j1 = total - j0 < kBinPairs ? total : j0 + kBinPairs;
    for (uint32_t jj = j0 + threadIdx.x; jj < j1; jj += kBinThreads) {
        const uint32_t i = jj / P, p = jj - i * P;
        uint32_t idx[8];
        float fr[3];
        pair_corners<P>(lv, in, i, p, idx, fr);
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t r = idx[c] / E - w0;
            if (r < bins) atomicAdd(&hist[r], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < bins; j += kBinThreads)
        if (hist[j] != 0u) atomicAdd(&counts[first + j], hist[j]);
}

// One workgroup: offsets[b] = counts[0] + ... + counts[b - 1] for b <= bins, cursors = offsets.
__global__ __launch_bounds__(1024) void train_bin_scan_kernel(const uint32_t* __restrict__ counts, uint32_t bins,
                                                              uint32_t* __restrict__ offsets,
                                                              uint32_t* __restrict__ cursors)
{
    __shared__ uint32_t part[1024];
    const uint32_t tid = threadIdx.x, per = (bins + 1023u) / 1024u;
    const uint32_t b0 = tid * per < bins ? tid * per : bins, b1 = bins - b0 < per ? bins : b0 + per;
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += counts[b];
    part[tid] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid > 0 ? part[tid - 1] : 0u;
    for (uint32_t b = b0; b < b1; b++) {
        offsets[b] = run;
        cursors[b] = run;
        run += counts[b];
    }
    if (tid == 1023u) offsets[bins] = part[1023];
}

// The count kernel's workgroups again: local counts, a range of each touched bin reserved with one
// global atomic, then each term (entry within its bin, F values) written at the range's next
// position (LDS counter).
template <int P, int L, int F>
__global__ __launch_bounds__(kBinThreads) void train_bin_scatter_kernel(TrainArgs t, BinMap bm,
                                                                        const float* __restrict__ in,
                                                                        const float* __restrict__ act, uint32_t n,
                                                                        uint32_t* __restrict__ cursors,
                                                                        uint32_t* __restrict__ term_r,
                                                                        float* __restrict__ term_v)
{
    extern __shared__ uint32_t lds[];
    constexpr uint32_t E = kBinFloats / F;
    const uint32_t k = blockIdx.y, l = bm.level[k], w0 = blockIdx.z * kBinWindow;
    const uint32_t level_bins = bm.first[k + 1] - bm.first[k];
    if (w0 >= level_bins) return;
    const uint32_t first = bm.first[k] + w0, bins = level_bins - w0 < kBinWindow ? level_bins - w0 : kBinWindow;
    uint32_t* hist = lds;
    uint32_t* base = lds + bins;
    const GridLevel& lv = t.net.level[l];
    for (uint32_t j = threadIdx.x; j < bins; j += kBinThreads) hist[j] = 0u;
    __syncthreads();
    const uint32_t total = n * (uint32_t)P, j0 = blockIdx.x * kBinPairs;
    const uint32_t j1 = total - j0 < kBinPairs ? total : j0 + kBinPairs;
    for (uint32_t jj = j0 + threadIdx.x; jj < j1; jj += kBinThreads) {
        const uint32_t i = jj / P, p = jj - i * P;
        uint32_t idx[8];
        float fr[3];
        pair_corners<P>(lv, in, i, p, idx, fr);
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const uint32_t r = idx[c] / E - w0;
            if (r < bins) atomicAdd(&hist[r], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < bins; j += kBinThreads) {
        const uint32_t c = hist[j];
        base[j] = c != 0u ? atomicAdd(&cursors[first + j], c) : 0u;
        hist[j] = 0u;
    }
    __syncthreads();
    const size_t S = t.act_stride;
    for (uint32_t jj = j0 + threadIdx.x; jj < j1; jj += kBinThreads) {
        const uint32_t i = jj / P, p = jj - i * P;
        uint32_t idx[8];
        float fr[3];
        pair_corners<P>(lv, in, i, p, idx, fr);
        float dx[F];
#pragma unroll
        for (int f = 0; f < F; f++) dx[f] = act[(size_t)(kDX + (p * L + l) * F + f) * S + i];
#pragma unroll
        for (uint32_t c = 0; c < 8; c++) {
            const uint32_t sl = idx[c] / E, r = sl - w0;
            if (r >= bins) continue;
            const uint32_t pos = base[r] + atomicAdd(&hist[r], 1u);
            const float w = corner_weight(c, fr);
            term_r[pos] = idx[c] - sl * E;
#pragma unroll
            for (int f = 0; f < F; f++) term_v[(size_t)pos * F + f] = w * dx[f];
        }
    }
}

// One workgroup per bin: the bin's terms added in LDS (kBinUnroll terms per thread loaded before
// their adds), the bin's entries stored whole (zeros included).
template <int F>
__global__ __launch_bounds__(kScatterThreads) void train_bin_accumulate_kernel(TrainArgs t, BinMap bm,
                                                                               const uint32_t* __restrict__ offsets,
                                                                               const uint32_t* __restrict__ term_r,
                                                                               const float* __restrict__ term_v,
                                                                               float* __restrict__ gtab)
{
    constexpr uint32_t E = kBinFloats / F;
    __shared__ float4 acc4[kBinFloats / 4];
    float* acc = reinterpret_cast<float*>(acc4);
    const uint32_t b = blockIdx.x;
    uint32_t k = 0;
    while (k + 1 < bm.count && b >= bm.first[k + 1]) k++;
    const GridLevel& lv = t.net.level[bm.level[k]];
    const uint32_t e0 = (b - bm.first[k]) * E;
    const uint32_t len = lv.size - e0 < E ? lv.size - e0 : E;  // a multiple of 8 entries
    for (uint32_t j = threadIdx.x; j < len * F / 4; j += kScatterThreads) acc4[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const uint32_t q0 = offsets[b], q1 = offsets[b + 1];
    constexpr uint32_t U = kBinUnroll, B = kScatterThreads;
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += U * B) {
        uint32_t r[U];
        float v[U][F];
#pragma unroll
        for (uint32_t h = 0; h < U; h++) {
            const uint32_t qq = q + h * B < q1 ? q + h * B : q;
            r[h] = term_r[qq];
#pragma unroll
            for (int f = 0; f < F; f++) v[h][f] = term_v[(size_t)qq * F + f];
        }
#pragma unroll
        for (uint32_t h = 0; h < U; h++) {
            if (q + h * B >= q1) break;
#pragma unroll
            for (int f = 0; f < F; f++) atomicAdd(&acc[r[h] * F + f], v[h][f]);
        }
    }
    __syncthreads();
    float4* out = reinterpret_cast<float4*>(gtab + ((size_t)lv.offset + e0) * F);
    for (uint32_t j = threadIdx.x; j < len * F / 4; j += kScatterThreads) out[j] = acc4[j];
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

// tiny-cuda-nn's AdamOptimizer step (public optimizers/adam.h, the reference's optimiser):
// MATRIX parameters (the MLP) get L2 regularisation added to the gradient and are updated every
// step, also with a zero gradient (their moments decay); the other parameters (the grid tables)
// are skipped where the gradient is exactly zero and get no L2 term. Each parameter counts its
// own updates t, and the bias correction is folded into the rate:
// w -= lr * sqrt(1 - beta2^t) / (1 - beta1^t) * m / (sqrt(v) + epsilon).
template <bool MATRIX>
__global__ __launch_bounds__(256) void train_adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                         float* __restrict__ m, float* __restrict__ v,
                                                         uint32_t* __restrict__ steps, uint64_t n, AdamArgs h)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    if (MATRIX)
        gi += h.l2_reg * w[i];
    else if (gi == 0.f)
        return;  // sparse update: entries no sample touched keep their value, moments and count
    const float mi = h.beta1 * m[i] + (1.f - h.beta1) * gi;
    const float vi = h.beta2 * v[i] + (1.f - h.beta2) * gi * gi;
    const uint32_t t = steps[i] + 1u;
    const float lr = h.lr * sqrtf(1.f - powf(h.beta2, (float)t)) / (1.f - powf(h.beta1, (float)t));
    m[i] = mi;
    v[i] = vi;
    steps[i] = t;
    w[i] = w[i] - lr * mi / (sqrtf(vi) + h.epsilon);
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

// Bin workspace: the bin counts, offsets and cursors, then the corner terms of a max batch
// (P * L * 8 per sample: entry within its slice, and F values).
constexpr uint32_t kMaxBins = 131072;  // bins of all binned levels (2^26 entries x 16 levels at 2 features)

size_t train_bin_bytes(uint32_t max_batch, uint32_t n_points, uint32_t n_levels, uint32_t n_features)
{
    const size_t terms = (size_t)train_act_stride(max_batch) * n_points * n_levels * 8u;
    return (size_t)3 * (kMaxBins + 1) * sizeof(uint32_t) + terms * (1 + n_features) * sizeof(float);
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
    if (a.n_levels > (uint32_t)kMaxLevels || a.n_features == 0) return hipErrorInvalidValue;
    // Levels of at least NBVH_BIN_MIN_SLICES slices (16 MB of gradient) take the binned path; the
    // others are scanned once per slice, split into parts so that each workgroup adds about 32K
    // corners if the batch's corners spread evenly over the level (a skewed batch only makes some
    // workgroups slower). Measured per 65,536-sample step of the 16 x 2 grid, binning levels of
    // >= 16 / 64 / 128 slices and slice scans only (2048 pairs per binning workgroup): 2^19 1.14 /
    // 1.06 / 1.06 / 1.06 ms, 2^20 1.37 / 1.40 / 1.32 / 1.32, 2^21 1.66 / 1.69 / 1.69 / 1.89, 2^22
    // 2.06 / 2.09 / 2.09 / 2.74, 2^23 2.60 / 2.62 / 2.62 / 4.03 (global atomics on the gradient,
    // the earlier form for the largest tables: 2^22 2.82, 2^23 3.83).
#ifndef NBVH_BIN_MIN_SLICES
#define NBVH_BIN_MIN_SLICES 128
#endif
    dev::SliceMap map{};
    dev::BinMap bins{};
    uint32_t max_bins = 0;
    for (uint32_t l = 0; l < a.n_levels; l++) {
        const uint32_t E = dev::kSliceFloats / a.n_features, size = a.level[l].size;
        const uint32_t slices = (size + E - 1) / E;
        const double corners = (double)n * a.n_points * 8.0 * std::min(E, size) / size;
        if (slices >= NBVH_BIN_MIN_SLICES) {
            const uint32_t EB = dev::kBinFloats / a.n_features, level_bins = (size + EB - 1) / EB;
            map.parts[l]     = 1;
            map.first[l + 1] = map.first[l];
            bins.level[bins.count]     = l;
            bins.first[bins.count + 1] = bins.first[bins.count] + level_bins;
            bins.count++;
            max_bins = std::max(max_bins, level_bins);
            continue;
        }
        map.parts[l]     = (uint32_t)std::min(64.0, std::max(1.0, std::ceil(corners / 32768.0)));
        map.first[l + 1] = map.first[l] + slices * map.parts[l];
    }
    const uint32_t n_bins = bins.first[bins.count];
    if (n_bins > kMaxBins) return hipErrorInvalidValue;
    // Term positions are 32-bit: a step bins fewer than 2^32 corner terms.
    if ((uint64_t)n * a.n_points * 8u * bins.count > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t windows = (max_bins + dev::kBinWindow - 1) / dev::kBinWindow;
    const uint32_t window_bins = std::min(max_bins, dev::kBinWindow);
    uint32_t* const counts  = reinterpret_cast<uint32_t*>(t.bins);
    uint32_t* const offsets = counts + (kMaxBins + 1);
    uint32_t* const cursors = offsets + (kMaxBins + 1);
    uint32_t* const term_r  = cursors + (kMaxBins + 1);
    float* const    term_v  = reinterpret_cast<float*>(term_r + (size_t)S * a.n_points * a.n_levels * 8u);
    const dim3 bin_grid((n * a.n_points + dev::kBinPairs - 1) / dev::kBinPairs, bins.count, windows);
    // Split slices add into the gradient with global atomics: their levels start from zero. Every
    // other entry is stored by its slice.
    for (uint32_t l = 0; l < a.n_levels; l++)
        if (map.parts[l] > 1) {
            const hipError_t e = train_zero(t.grad_tables + (size_t)a.level[l].offset * a.n_features,
                                            (uint64_t)a.level[l].size * a.n_features, stream);
            if (e != hipSuccess) return e;
        }
    if (n_bins > 0) {
        const hipError_t e = train_zero(reinterpret_cast<float*>(counts), n_bins, stream);
        if (e != hipSuccess) return e;
    }
    bool shape = false;
#define X(P_, L_, F_)                                                                                                \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_) {                                                \
        shape = true;                                                                                                \
        hipLaunchKernelGGL(dev::train_transpose_kernel, grid_mlp, dim3(256), 0, stream, t.mlp, wt);                  \
        hipLaunchKernelGGL((dev::train_forward_kernel<P_, L_, F_>), grid, block, 0, stream, t, t.mlp, in, act, n);    \
        hipLaunchKernelGGL((dev::train_backward_kernel<P_, L_, F_>), grid, block, 0, stream, t, (const float*)wt, in, \
                           targets, act, n);                                                                         \
        if (map.first[L_] > 0)                                                                                       \
            hipLaunchKernelGGL((dev::train_scatter_kernel<P_, L_, F_>), dim3(map.first[L_]), dim3(dev::kScatterThreads), \
                               0, stream, t, map, in, (const float*)act, t.grad_tables, n);                          \
        if (n_bins > 0) {                                                                                            \
            hipLaunchKernelGGL((dev::train_bin_count_kernel<P_, F_>), bin_grid, dim3(dev::kBinThreads),                \
                               window_bins * 4, stream, t, bins, in, n, counts);                                     \
            hipLaunchKernelGGL(dev::train_bin_scan_kernel, dim3(1), dim3(1024), 0, stream, (const uint32_t*)counts,  \
                               n_bins, offsets, cursors);                                                           \
            hipLaunchKernelGGL((dev::train_bin_scatter_kernel<P_, L_, F_>), bin_grid, dim3(dev::kBinThreads),          \
                               window_bins * 8, stream, t, bins, in, (const float*)act, n, cursors, term_r, term_v);  \
            hipLaunchKernelGGL((dev::train_bin_accumulate_kernel<F_>), dim3(n_bins), dim3(dev::kScatterThreads), 0,   \
                               stream, t, bins, (const uint32_t*)offsets, (const uint32_t*)term_r,                  \
                               (const float*)term_v, t.grad_tables);                                                \
        }                                                                                                            \
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

hipError_t train_adam(float* w, const float* g, float* m, float* v, uint32_t* steps, uint64_t n, bool matrix,
                      const AdamArgs& h, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (matrix)
        hipLaunchKernelGGL(dev::train_adam_kernel<true>, grid, dim3(256), 0, stream, w, g, m, v, steps, n, h);
    else
        hipLaunchKernelGGL(dev::train_adam_kernel<false>, grid, dim3(256), 0, stream, w, g, m, v, steps, n, h);
    return hipGetLastError();
}

}  // namespace nbvh
