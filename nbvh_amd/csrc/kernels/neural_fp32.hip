// neural_fp32.hip -- the plain FP32 neural query (SIMT FMA MLP, FP32 hash tables) and the FP32
// network inference. One lane evaluates one segment; MLP weights are wave-uniform (scalar loads),
// activations live in registers with one 64x64 LDS buffer per wave between layers.
//
// Arithmetic order is that of oracle/nbvh_oracle.cpp (FP32 mode): the FP32 path is bit-exact
// with it.
#include <hip/hip_runtime.h>

#include "../device/wave_query.h"
#include "launch.h"

namespace nbvh {
namespace dev {

// MLP with 4 hidden layers of 64 (5 matrices). Weight layout of a.mlp_f32 for this kernel:
// layers 1, 3, 5 row-major (out x in), layers 2, 4 transposed (in x out), see
// host/api.cpp:fp32_kernel_layout. Accumulation: acc = fmaf(W[j][k], x[k], acc), k ascending.
// Weights are read through the constant address space: they are read-only for the whole launch,
// so wave-uniform reads become scalar loads into SGPRs (v_fma with an SGPR operand) instead of
// one VGPR per weight.
typedef const float __attribute__((address_space(4))) cfloat;

template <int K0>
__device__ __forceinline__ void mlp_f32(const float* W_generic, const float* x, float* h_lds, float z[8])
{
    const cfloat* __restrict__ W = (const cfloat*)W_generic;
    const uint32_t lane = threadIdx.x & 63u;
    // Layer 1: j rolled, inputs in registers, output to LDS.
    for (int j = 0; j < 64; j++) {
        const cfloat* w = W +j * K0;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < K0; k++) acc = fmaf(w[k], x[k], acc);
        h_lds[j * 64 + lane] = acc > 0.f ? acc : 0.f;
    }
    W += 64 * K0;
    float r[64];
    // Layer 2: k rolled from LDS, outputs in registers.
#pragma unroll
    for (int j = 0; j < 64; j++) r[j] = 0.f;
    for (int k = 0; k < 64; k++) {
        const float hk = h_lds[k * 64 + lane];
        const cfloat* w = W +k * 64;
#pragma unroll
        for (int j = 0; j < 64; j++) r[j] = fmaf(w[j], hk, r[j]);
    }
#pragma unroll
    for (int j = 0; j < 64; j++) r[j] = r[j] > 0.f ? r[j] : 0.f;
    W += 64 * 64;
    // Layer 3: j rolled, inputs in registers, output to LDS.
    for (int j = 0; j < 64; j++) {
        const cfloat* w = W +j * 64;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 64; k++) acc = fmaf(w[k], r[k], acc);
        h_lds[j * 64 + lane] = acc > 0.f ? acc : 0.f;
    }
    W += 64 * 64;
    // Layer 4: k rolled from LDS.
#pragma unroll
    for (int j = 0; j < 64; j++) r[j] = 0.f;
    for (int k = 0; k < 64; k++) {
        const float hk = h_lds[k * 64 + lane];
        const cfloat* w = W +k * 64;
#pragma unroll
        for (int j = 0; j < 64; j++) r[j] = fmaf(w[j], hk, r[j]);
    }
#pragma unroll
    for (int j = 0; j < 64; j++) r[j] = r[j] > 0.f ? r[j] : 0.f;
    W += 64 * 64;
    // Output layer (8 x 64), row-major.
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 64; k++) acc = fmaf(W[j * 64 + k], r[k], acc);
        z[j] = acc;
    }
}

template <int P, int L, int F, bool PARTS>
// Minimum waves per SIMD for the register allocation. Measured: forcing 2 halves the FP32
// query's throughput (the allocator then starves the scalar weight prefetch), so 1 (no bound).
#ifndef NBVH_FP32_WAVES_PER_EU
#define NBVH_FP32_WAVES_PER_EU 1
#endif
__global__ __launch_bounds__(64, NBVH_FP32_WAVES_PER_EU) void neural_query_fp32_kernel(NetArgs a, RayBatch b,
                                                               uint32_t* __restrict__ queue)
{
    constexpr int K0 = P * L * F;
    __shared__ uint32_t stack_lds[kNeuralStackSize * 64];
    __shared__ float h_lds[64 * 64];
    const uint32_t lane = threadIdx.x;
    LdsStack st{stack_lds + lane, 0};
    Slot s;
    s.state = SLOT_EMPTY;
    RayQueue q;
    q.init(queue, batch_count(b), a.image_width, a.queue_bands);
    while (fill_slots<PARTS>(a, GlobalNodes{a.nodes}, s, st, b, q)) {
        float x[K0];
        if (s.state == SLOT_SEG) {
            float pts[3 * P];
            encode_segment(a, s.ray, s.seg_mint, s.seg_maxt, pts, &s.rng);
            grid_features<P, L, F, float>(a, a.tables_f32, pts, x);
        } else {
#pragma unroll
            for (int k = 0; k < K0; k++) x[k] = 0.f;
        }
        float z[8];
        mlp_f32<K0>(a.mlp_f32, x, h_lds, z);
        if (s.state == SLOT_SEG) consume_segment<PARTS>(a, s, z, b);
    }
}

// tcnn::Network::inference replacement, FP32: in[n x 3P] -> out[n x 8] (sigmoid), logits.
template <int P, int L, int F>
__global__ __launch_bounds__(64) void network_fp32_kernel(NetArgs a, const float* __restrict__ in, uint32_t n,
                                                          float* __restrict__ out, float* __restrict__ logits)
{
    constexpr int K0 = P * L * F;
    __shared__ float h_lds[64 * 64];
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    float x[K0];
    if (i < n) {
        float pts[3 * P];
#pragma unroll
        for (int c = 0; c < 3 * P; c++) pts[c] = in[(size_t)i * (3 * P) + c];
        grid_features<P, L, F, float>(a, a.tables_f32, pts, x);
    } else {
#pragma unroll
        for (int k = 0; k < K0; k++) x[k] = 0.f;
    }
    float z[8];
    mlp_f32<K0>(a.mlp_f32, x, h_lds, z);
    if (i < n) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (out) out[(size_t)i * 8 + j] = sigmoidf_cw(z[j]);
            if (logits) logits[(size_t)i * 8 + j] = z[j];
        }
    }
}

}  // namespace dev

template <int P, int L, int F, bool PARTS>
static hipError_t launch_fp32_query_k(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                                      hipStream_t stream)
{
    auto k = dev::neural_query_fp32_kernel<P, L, F, PARTS>;
    static const uint32_t bpc = [&] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, 64, 0) != hipSuccess || b <= 0) {
            (void)hipGetLastError();
            b = 1;
        }
        return (uint32_t)b;
    }();
    uint32_t blocks = (rb.n + 63) / 64;
    const uint32_t cap = a.blocks_per_cu != 0u && a.blocks_per_cu < bpc ? a.blocks_per_cu : bpc;
    blocks = blocks < n_cus * cap ? blocks : n_cus * cap;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64), 0, stream, a, rb, queue);
    return hipGetLastError();
}

template <int P, int L, int F>
static hipError_t launch_fp32_query(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                                    hipStream_t stream)
{
    return rb.parts ? launch_fp32_query_k<P, L, F, true>(a, rb, queue, n_cus, stream)
                    : launch_fp32_query_k<P, L, F, false>(a, rb, queue, n_cus, stream);
}

template <int P, int L, int F>
static hipError_t launch_fp32_network(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits,
                                      hipStream_t stream)
{
    hipLaunchKernelGGL((dev::network_fp32_kernel<P, L, F>), dim3((n + 63) / 64), dim3(64), 0, stream, a, in, n, out,
                       logits);
    return hipGetLastError();
}

#define NBVH_FP32_SHAPES(X) X(3, 16, 2) X(3, 8, 4) X(3, 2, 2) X(3, 4, 2)

hipError_t query_fp32(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream,
                      bool* supported)
{
    *supported = true;
#define X(P_, L_, F_)                                                                                    \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_ && a.n_hidden_layers == 4)            \
        return launch_fp32_query<P_, L_, F_>(a, rb, queue, n_cus, stream);
    NBVH_FP32_SHAPES(X)
#undef X
    *supported = false;
    return hipSuccess;
}

hipError_t network_fp32(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits, hipStream_t stream,
                        bool* supported)
{
    *supported = true;
#define X(P_, L_, F_)                                                                                    \
    if (a.n_points == P_ && a.n_levels == L_ && a.n_features == F_ && a.n_hidden_layers == 4)            \
        return launch_fp32_network<P_, L_, F_>(a, in, n, out, logits, stream);
    NBVH_FP32_SHAPES(X)
#undef X
    *supported = false;
    return hipSuccess;
}

}  // namespace nbvh
