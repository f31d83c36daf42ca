// wavefront.hip -- sorted wavefront rounds of the neural query (NBVH_QUERY_WAVEFRONT).
//
// The persistent kernels evaluate the segments of whatever 64 rays a wave holds; for incoherent
// rays those segments lie all over the scene and the hash-grid gathers miss in L1 and L2 (about
// 3.5x the per-segment gather cost of coherent primary rays). Here a query runs in rounds, like
// the reference's step loop (NeuralPathModule::neural_bvh_traversal,
// src/cuda/neural/path/path_module.cu:54-137), but the round's segments are evaluated in the
// order of their midpoints' Morton codes, so the 64 segments of a wave lie close together and
// share coarse grid cells (measured on 4M incoherent rays: network time 28.5 ms unsorted,
// 16.6 ms grouped by leaf, 14.5 ms in Morton order):
//   traverse  every active ray advances to its next accepted leaf (one thread per ray, global
//             SoA stack, high occupancy) and appends the segment with its sort key;
//   sort      segment indices by key (radix sort over 30 bits);
//   encode    the segments in key order (9 floats each);
//   network   the network kernels on the encoded segments (the same MLP code as the query);
//   consume   apply_segment on each ray's segment (the same decision code as the persistent
//             scheduler); finished rays write their hit record, the others stay active.
// Every ray sees the same segments in the same order with the same arithmetic, so the hit
// records are bit-identical to the persistent kernels'.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "../device/wave_query.h"
#include "launch.h"

namespace nbvh {
namespace dev {

struct WfRay {
    HitState h;
    uint32_t stack_n;
    uint32_t pad[3];
};
static_assert(sizeof(WfRay) == 64, "wavefront ray state is 64 B");

// Traversal stack of ray g in global memory, entry e at col[e * stride] (coalesced across rays).
struct GlobalStack {
    uint32_t* col;
    uint32_t stride;
    int n;
    __device__ __forceinline__ void push(uint32_t v)
    {
        col[(size_t)n * stride] = v;
        n++;
    }
    __device__ __forceinline__ uint32_t pop()
    {
        n--;
        return col[(size_t)n * stride];
    }
    __device__ __forceinline__ bool empty() const { return n == 0; }
    __device__ __forceinline__ int size() const { return n; }
};

__device__ __forceinline__ void wf_locate(const RayBatch& b, uint32_t g, uint32_t& part, uint32_t& local)
{
    if (b.parts) {
        part  = batch_part(b, g);
        local = g - b.parts[part].begin;
    } else {
        part  = 0;
        local = g;
    }
}

__device__ __forceinline__ bool wf_any_hit(const NetArgs& a, const RayBatch& b, uint32_t part)
{
    return b.parts ? b.parts[part].any_hit != 0u : a.any_hit != 0u;
}

__global__ __launch_bounds__(256) void wf_init_kernel(RayBatch b, WfRay* __restrict__ st, uint32_t* __restrict__ stacks,
                                                      uint32_t* __restrict__ active)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b.n) return;
    uint32_t part, local;
    wf_locate(b, g, part, local);
    const Ray r = load_ray(b.parts ? b.parts[part].rays : b.rays, local);
    WfRay w;
    hit_state_init(w.h, r.tmax);
    w.stack_n = 1;
    w.pad[0] = w.pad[1] = w.pad[2] = 0;
    st[g]     = w;
    stacks[g] = 0u;  // root
    active[g] = g;
}

__global__ __launch_bounds__(256) void wf_traverse_kernel(NetArgs a, RayBatch b, WfRay* __restrict__ st,
                                                          uint32_t* __restrict__ stacks,
                                                          const uint32_t* __restrict__ active, uint32_t n_active,
                                                          WfSeg* __restrict__ segs, uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals, uint32_t* __restrict__ counter)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_active) return;
    const uint32_t g = active[i];
    uint32_t part, local;
    wf_locate(b, g, part, local);
    const Ray r = load_ray(b.parts ? b.parts[part].rays : b.rays, local);
    WfRay w = st[g];
    GlobalStack stk{stacks + g, b.n, (int)w.stack_n};
    int res = 0;
    float mint = 0.f, maxt = 0.f;
    uint32_t node = 0;
    if (!stk.empty()) res = nbvh_next_leaf(GlobalNodes{a.nodes}, a.lod, r, stk, w.h.best_t, mint, maxt, node);
    if (res == 1 && w.h.segments != a.max_steps + 1u) {
        const uint32_t k = atomicAdd(counter, 1u);
        segs[k] = WfSeg{g, node, mint, maxt};
        // Sort key: Morton code of the segment's midpoint in the encoding box.
        const float tm = 0.5f * (mint + maxt);
        keys[k] = morton30((r.ox + tm * r.dx - a.enc_center[0]) / a.enc_extent[0] + 0.5f,
                           (r.oy + tm * r.dy - a.enc_center[1]) / a.enc_extent[1] + 0.5f,
                           (r.oz + tm * r.dz - a.enc_center[2]) / a.enc_extent[2] + 0.5f);
        vals[k] = k;
        st[g].stack_n = (uint32_t)stk.n;
        return;
    }
    if (res == 1) w.h.flags |= NBVH_FLAG_STEP_CAP;
    if (res < 0) w.h.flags |= NBVH_FLAG_STACK_OVERFLOW;
    const bool any = wf_any_hit(a, b, part);
    store_hit(b.parts ? b.parts[part].hits : b.hits, local, make_hit(w.h, any));
}

template <int P>
__global__ __launch_bounds__(256) void wf_encode_kernel(NetArgs a, RayBatch b, const WfSeg* __restrict__ segs,
                                                        const uint32_t* __restrict__ order, uint32_t m,
                                                        float* __restrict__ in)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const WfSeg sg = segs[order[j]];
    uint32_t part, local;
    wf_locate(b, sg.ray, part, local);
    const Ray r = load_ray(b.parts ? b.parts[part].rays : b.rays, local);
    float pts[3 * P];
    encode_segment(a, r, sg.mint, sg.maxt, pts);
#pragma unroll
    for (int c = 0; c < 3 * P; c++) in[(size_t)j * (3 * P) + c] = pts[c];
}

__global__ __launch_bounds__(256) void wf_consume_kernel(NetArgs a, RayBatch b, const WfSeg* __restrict__ segs,
                                                         const uint32_t* __restrict__ order,
                                                         const float* __restrict__ logits, uint32_t m,
                                                         WfRay* __restrict__ st, uint32_t* __restrict__ next,
                                                         uint32_t* __restrict__ next_count)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const WfSeg sg = segs[order[j]];
    const uint32_t g = sg.ray;
    uint32_t part, local;
    wf_locate(b, g, part, local);
    const bool any = wf_any_hit(a, b, part);
    float z[8];
#pragma unroll
    for (int k = 0; k < 8; k++) z[k] = logits[(size_t)j * 8 + k];
    HitState h = st[g].h;
    Ray r{};
    if (a.depth_mode == NBVH_DEPTH_WORLD_T) r = load_ray(b.parts ? b.parts[part].rays : b.rays, local);
    if (apply_segment(a, h, any, r, sg.mint, sg.maxt, sg.node, z)) {
        store_hit(b.parts ? b.parts[part].hits : b.hits, local, make_hit(h, any));
        return;
    }
    st[g].h = h;
    next[atomicAdd(next_count, 1u)] = g;
}

}  // namespace dev

hipError_t wf_init(const RayBatch& b, WfRayState* st, uint32_t* stacks, uint32_t* active, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_init_kernel, dim3((b.n + 255) / 256), dim3(256), 0, stream, b,
                       reinterpret_cast<dev::WfRay*>(st), stacks, active);
    return hipGetLastError();
}

hipError_t wf_traverse(const NetArgs& a, const RayBatch& b, WfRayState* st, uint32_t* stacks, const uint32_t* active,
                       uint32_t n_active, WfSeg* segs, uint32_t* keys, uint32_t* vals, uint32_t* counter,
                       hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_traverse_kernel, dim3((n_active + 255) / 256), dim3(256), 0, stream, a, b,
                       reinterpret_cast<dev::WfRay*>(st), stacks, active, n_active, segs, keys, vals, counter);
    return hipGetLastError();
}

hipError_t wf_encode(const NetArgs& a, const RayBatch& b, const WfSeg* segs, const uint32_t* order, uint32_t m,
                     float* in, hipStream_t stream)
{
    switch (a.n_points) {
    case 1: hipLaunchKernelGGL(dev::wf_encode_kernel<1>, dim3((m + 255) / 256), dim3(256), 0, stream, a, b, segs, order, m, in); break;
    case 2: hipLaunchKernelGGL(dev::wf_encode_kernel<2>, dim3((m + 255) / 256), dim3(256), 0, stream, a, b, segs, order, m, in); break;
    case 3: hipLaunchKernelGGL(dev::wf_encode_kernel<3>, dim3((m + 255) / 256), dim3(256), 0, stream, a, b, segs, order, m, in); break;
    case 4: hipLaunchKernelGGL(dev::wf_encode_kernel<4>, dim3((m + 255) / 256), dim3(256), 0, stream, a, b, segs, order, m, in); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t wf_consume(const NetArgs& a, const RayBatch& b, const WfSeg* segs, const uint32_t* order,
                      const float* logits, uint32_t m, WfRayState* st, uint32_t* next, uint32_t* next_count,
                      hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_consume_kernel, dim3((m + 255) / 256), dim3(256), 0, stream, a, b, segs, order, logits,
                       m, reinterpret_cast<dev::WfRay*>(st), next, next_count);
    return hipGetLastError();
}

size_t wf_sort_scratch_bytes(uint32_t n)
{
    size_t tmp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
    return tmp;
}

hipError_t wf_sort(void* tmp, size_t tmp_bytes, const uint32_t* keys, uint32_t* keys_out, const uint32_t* vals,
                   uint32_t* order, uint32_t m, int end_bit, hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_out, vals, order, (int)m, 0, end_bit, stream);
}

}  // namespace nbvh
