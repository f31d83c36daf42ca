// reorder.hip -- coherence order for incoherent ray batches.
//
// Rays are traced in the order of a 40-bit key: part of a multi-part launch (6 bits), direction
// octant (3 bits) and the 30-bit Morton code of the origin quantised over the scene's encoding
// box (10 bits per axis). Neighbouring claims then start in neighbouring leaves and gather from
// neighbouring grid cells, which the L1/L2 serve from lines their neighbours just brought in.
// The permutation is consumed by the query kernels through RayBatch::order; hits are still
// written at each ray's own index, so the result is the same as in array order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "../device/common.h"
#include "launch.h"

namespace nbvh {
namespace dev {

__global__ __launch_bounds__(256) void ray_keys_kernel(RayBatch b, float3 lo, float3 inv_ext,
                                                       unsigned long long* __restrict__ keys,
                                                       uint32_t* __restrict__ ids)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= b.n) return;
    uint32_t part = 0, i = g;
    if (b.parts) {
        part = batch_part(b, g);
        i    = g - b.parts[part].begin;
    }
    const Ray r = load_ray(b.parts ? b.parts[part].rays : b.rays, i);
    const float q[3] = {(r.ox - lo.x) * inv_ext.x, (r.oy - lo.y) * inv_ext.y, (r.oz - lo.z) * inv_ext.z};
    const uint32_t m = morton30(q[0], q[1], q[2]);
    const uint32_t oct = (r.dx < 0.f ? 1u : 0u) | (r.dy < 0.f ? 2u : 0u) | (r.dz < 0.f ? 4u : 0u);
    keys[g] = ((unsigned long long)part << 33) | ((unsigned long long)oct << 30) | m;
    ids[g]  = g;
}

}  // namespace dev

// Scratch layout: keys in/out (8 B each), ids in (4 B), then the radix sort's temporary storage.
size_t ray_order_scratch_bytes(uint32_t n)
{
    size_t tmp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const unsigned long long*)nullptr,
                                             (unsigned long long*)nullptr, (const uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (int)n, 0, 39);
    const size_t a = ((size_t)n * 20 + 255) / 256 * 256;
    return a + tmp;
}

hipError_t ray_order(const RayBatch& rb, const float enc_min[3], const float enc_ext[3], uint32_t* order,
                     void* scratch, size_t scratch_bytes, hipStream_t stream)
{
    const uint32_t n = rb.n;
    unsigned long long* k_in  = static_cast<unsigned long long*>(scratch);
    unsigned long long* k_out = k_in + n;
    uint32_t* ids             = reinterpret_cast<uint32_t*>(k_out + n);
    const size_t a            = ((size_t)n * 20 + 255) / 256 * 256;
    void* tmp                 = static_cast<char*>(scratch) + a;
    size_t tmp_bytes          = scratch_bytes - a;
    const float3 lo  = make_float3(enc_min[0], enc_min[1], enc_min[2]);
    const float3 inv = make_float3(enc_ext[0] > 0.f ? 1.f / enc_ext[0] : 0.f, enc_ext[1] > 0.f ? 1.f / enc_ext[1] : 0.f,
                                   enc_ext[2] > 0.f ? 1.f / enc_ext[2] : 0.f);
    hipLaunchKernelGGL(dev::ray_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, rb, lo, inv, k_in, ids);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, ids, order, (int)n, 0, 39, stream);
}

}  // namespace nbvh
