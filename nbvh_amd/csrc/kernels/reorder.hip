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
    if (g >= batch_count(b)) {  // beyond a device-side count: sorted after every live ray
        keys[g] = ~0ull;
        ids[g]  = g;
        return;
    }
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

// Bucket order (the form that can be recorded into a HIP graph: kernels only, no library
// memsets): 2^16 buckets of (direction octant, top 13 bits of the origin's Morton code), a
// counting pass, an exclusive scan and a scatter; rays of one bucket keep no particular order.
constexpr uint32_t kOrderBuckets = 1u << 16;

__device__ __forceinline__ uint32_t ray_bucket(const Ray& r, float3 lo, float3 inv_ext)
{
    const uint32_t m = morton30((r.ox - lo.x) * inv_ext.x, (r.oy - lo.y) * inv_ext.y, (r.oz - lo.z) * inv_ext.z);
    const uint32_t oct = (r.dx < 0.f ? 1u : 0u) | (r.dy < 0.f ? 2u : 0u) | (r.dz < 0.f ? 4u : 0u);
    return (oct << 13) | (m >> 17);
}

__global__ __launch_bounds__(256) void bucket_count_kernel(RayBatch b, float3 lo, float3 inv_ext,
                                                           uint32_t* __restrict__ bins,
                                                           uint32_t* __restrict__ bucket)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= batch_count(b)) return;
    uint32_t part = 0, i = g;
    if (b.parts) {
        part = batch_part(b, g);
        i    = g - b.parts[part].begin;
    }
    const uint32_t k = ray_bucket(load_ray(b.parts ? b.parts[part].rays : b.rays, i), lo, inv_ext);
    bucket[g] = k;
    atomicAdd(bins + k, 1u);
}

// Exclusive scan of the bucket counts in place: one workgroup of 1024 threads, 64 buckets each.
__global__ __launch_bounds__(1024) void bucket_scan_kernel(uint32_t* __restrict__ bins)
{
    constexpr uint32_t kPer = kOrderBuckets / 1024;
    __shared__ uint32_t part[1024];
    uint32_t* own = bins + threadIdx.x * kPer;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kPer; k++) sum += own[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;  // exclusive prefix of this thread's buckets
    for (uint32_t k = 0; k < kPer; k++) {
        const uint32_t c = own[k];
        own[k] = run;
        run += c;
    }
}

__global__ __launch_bounds__(256) void bucket_scatter_kernel(RayBatch b, const uint32_t* __restrict__ bucket,
                                                             uint32_t* __restrict__ cursor,
                                                             uint32_t* __restrict__ order)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= batch_count(b)) return;
    order[atomicAdd(cursor + bucket[g], 1u)] = g;
}

}  // namespace dev

size_t ray_bucket_scratch_bytes(uint32_t n)
{
    return (size_t)dev::kOrderBuckets * sizeof(uint32_t) + ((size_t)n * sizeof(uint32_t) + 255) / 256 * 256;
}

hipError_t ray_bucket_order(const RayBatch& rb, const float enc_min[3], const float enc_ext[3], uint32_t* order,
                            void* scratch, hipStream_t stream)
{
    const uint32_t n = rb.n;
    uint32_t* bins   = static_cast<uint32_t*>(scratch);
    uint32_t* bucket = bins + dev::kOrderBuckets;
    const float3 lo  = make_float3(enc_min[0], enc_min[1], enc_min[2]);
    const float3 inv = make_float3(enc_ext[0] > 0.f ? 1.f / enc_ext[0] : 0.f, enc_ext[1] > 0.f ? 1.f / enc_ext[1] : 0.f,
                                   enc_ext[2] > 0.f ? 1.f / enc_ext[2] : 0.f);
    hipError_t e = zero_u32(bins, dev::kOrderBuckets, stream);
    if (e != hipSuccess) return e;
    const dim3 grid((n + 255) / 256);
    hipLaunchKernelGGL(dev::bucket_count_kernel, grid, dim3(256), 0, stream, rb, lo, inv, bins, bucket);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(dev::bucket_scan_kernel, dim3(1), dim3(1024), 0, stream, bins);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(dev::bucket_scatter_kernel, grid, dim3(256), 0, stream, rb, bucket, bins, order);
    return hipGetLastError();
}

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
