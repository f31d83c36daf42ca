// basic.hip -- primary ray generation, counter reset and the triangle BVH fallback.
#include <hip/hip_runtime.h>

#include "../device/common.h"
#include "../device/sampling.h"
#include "../device/triangles.h"
#include "launch.h"

namespace nbvh {
namespace dev {

// camera_raygen_and_intersect_neural_bvh's ray (neural_bvh_module_common.cuh:230-236) and
// camera_raygen (neural_bvh_renderer_kernels.cuh:74-79).
__global__ void raygen_kernel(nbvh_camera cam, uint32_t w, uint32_t h, uint32_t sample, uint32_t frame,
                              nbvh_ray* __restrict__ rays)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t px = i % w, py = i / w;
    uint32_t st = lcg_seed(sample, frame, i);
    const float xf = ((float)px + lcg_float(st)) / (float)w;
    const float yf = ((float)py + lcg_float(st)) / (float)h;
    const float dx = (xf * cam.du[0] + yf * cam.dv[0]) + cam.dir_top_left[0];
    const float dy = (xf * cam.du[1] + yf * cam.dv[1]) + cam.dir_top_left[1];
    const float dz = (xf * cam.du[2] + yf * cam.dv[2]) + cam.dir_top_left[2];
    const float inv = 1.f / sqrtf((dx * dx + dy * dy) + dz * dz);
    float4* p = reinterpret_cast<float4*>(rays + i);
    p[0] = make_float4(cam.origin[0], cam.origin[1], cam.origin[2], 1e-4f);
    p[1] = make_float4(dx * inv, dy * inv, dz * inv, FLT_MAX);
}

// One diffuse bounce (path_inference_kernels.cu:380-393): Lambertian direction about the shading
// normal (sample_lambertian_dir + cos_sample_hemisphere, material.cuh:101-120,193-201, frame from
// ortho_basis, utils/common.cuh:21-35), origin offset by normal_epsilon_offset = 1e-4 along it
// (path_module.h:220). The shading normal is the hit's predicted normal (neural) or the
// triangle's geometric normal, turned toward the incoming ray. Random numbers: the reference's
// LCG seeded like raygen with frame = bounce + 1. Rays without a hit come out inactive (tmax = -1).
__global__ void scatter_diffuse_kernel(TriArgs tri, const nbvh_ray* __restrict__ rays, const nbvh_hit* __restrict__ hits,
                                       uint32_t n, uint32_t sample, uint32_t bounce, nbvh_ray* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Ray r = load_ray(rays, i);
    const nbvh_hit h = hits[i];
    float4* o = reinterpret_cast<float4*>(out + i);
    float nrm[3];
    if (!shading_normal(tri, h, r.dx, r.dy, r.dz, nrm)) {
        o[0] = make_float4(0.f, 0.f, 0.f, 0.f);
        o[1] = make_float4(0.f, 0.f, 1.f, -1.f);
        return;
    }
    uint32_t st = lcg_seed(sample, bounce + 1u, i);
    const float u0 = lcg_float(st), u1 = lcg_float(st);
    float wi[3];
    (void)cosine_direction(u0, u1, nrm, wi);
    float p[3] = {r.ox + h.t * r.dx, r.oy + h.t * r.dy, r.oz + h.t * r.dz};
    for (int c = 0; c < 3; c++) p[c] = p[c] + nrm[c] * 1e-4f;
    o[0] = make_float4(p[0], p[1], p[2], 1e-4f);
    o[1] = make_float4(wi[0], wi[1], wi[2], FLT_MAX);
}

// Hybrid query helpers. clip: the neural pass only looks in front of the triangle hit
// (tmax = min(tmax, t_tri)); merge: the neural hit replaces the triangle record when there is one
// (it is then closer, or for any-hit queries the ray was not occluded by triangles).
// live: optional device-side count (at most n) of the rays to process.
__device__ __forceinline__ uint32_t live_count(const uint32_t* live, uint32_t n)
{
    return live && *live < n ? *live : n;
}

__global__ void hybrid_clip_kernel(const nbvh_ray* __restrict__ rays, const nbvh_hit* __restrict__ tri, uint32_t n,
                                   const uint32_t* __restrict__ live, nbvh_ray* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= live_count(live, n)) return;
    Ray r = load_ray(rays, i);
    const nbvh_hit h = tri[i];
    if ((h.info & 0xFFu) != 0u && h.t < r.tmax) r.tmax = h.t;
    float4* o = reinterpret_cast<float4*>(out + i);
    o[0] = make_float4(r.ox, r.oy, r.oz, r.tmin);
    o[1] = make_float4(r.dx, r.dy, r.dz, r.tmax);
}

__global__ void hybrid_merge_kernel(nbvh_hit* __restrict__ hits, const nbvh_hit* __restrict__ neural, uint32_t n,
                                    const uint32_t* __restrict__ live, uint32_t any_hit)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= live_count(live, n)) return;
    const nbvh_hit nh = neural[i];
    const uint32_t tri_kind = hits[i].info & 0xFFu;
    if ((nh.info & 0xFFu) != 0u && !(any_hit && tri_kind != 0u)) store_hit(hits, i, nh);
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <bool PARTS>
__global__ __launch_bounds__(256) void tri_query_kernel(TriArgs a, RayBatch b)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= batch_count(b)) return;
    uint32_t part, i;
    batch_locate<PARTS>(b, batch_ray(b, g), part, i);
    const Ray ray = load_ray(batch_rays<PARTS>(b, part), i);
    const TriResult r = tri_trace(a, ray, batch_any_hit<PARTS>(b, part, a.any_hit));
    nbvh_hit h;
    h.t       = r.t;
    h.prim_id = r.prim;
    h.node_id = 0xFFFFFFFFu;
    h.info    = (r.found ? (uint32_t)NBVH_HIT_TRIANGLE : 0u) | (r.flags << 8);
    h.normal_or_uv[0] = r.u;
    h.normal_or_uv[1] = r.v;
    h.normal_or_uv[2] = 0.f;
    h.logit     = 0.f;
    h.albedo[0] = h.albedo[1] = h.albedo[2] = 0.f;
    h.reserved  = 0.f;
    store_hit(batch_hits<PARTS>(b, part), i, h);
}

// Network targets of n segments, ray i clipped to [mint[i], maxt[i]) (segment_targets).
__global__ __launch_bounds__(256) void segment_labels_kernel(TriArgs a, const nbvh_ray* __restrict__ rays,
                                                             const float* __restrict__ mint,
                                                             const float* __restrict__ maxt, uint32_t n,
                                                             float* __restrict__ targets)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float out[8];
    segment_targets(a, load_ray(rays, i), mint[i], maxt[i], out);
    for (int j = 0; j < 8; j++) targets[(size_t)i * 8 + j] = out[j];
}

}  // namespace dev

hipError_t raygen(const nbvh_camera& cam, uint32_t w, uint32_t h, uint32_t sample, uint32_t frame, nbvh_ray* rays,
                  hipStream_t stream)
{
    const uint32_t n = w * h;
    hipLaunchKernelGGL(dev::raygen_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, cam, w, h, sample, frame, rays);
    return hipGetLastError();
}

hipError_t zero_u32(uint32_t* p, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::zero_u32_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t scatter_diffuse(const TriArgs& tri, const nbvh_ray* rays, const nbvh_hit* hits, uint32_t n, uint32_t sample,
                           uint32_t bounce, nbvh_ray* out, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::scatter_diffuse_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, tri, rays, hits, n,
                       sample, bounce, out);
    return hipGetLastError();
}

hipError_t segment_labels(const TriArgs& a, const nbvh_ray* rays, const float* mint, const float* maxt, uint32_t n,
                          float* targets, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::segment_labels_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, rays, mint, maxt, n,
                       targets);
    return hipGetLastError();
}

hipError_t hybrid_clip(const nbvh_ray* rays, const nbvh_hit* tri, uint32_t n, const uint32_t* live, nbvh_ray* out,
                       hipStream_t stream)
{
    hipLaunchKernelGGL(dev::hybrid_clip_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, rays, tri, n, live, out);
    return hipGetLastError();
}

hipError_t hybrid_merge(nbvh_hit* hits, const nbvh_hit* neural, uint32_t n, const uint32_t* live, bool any_hit,
                        hipStream_t stream)
{
    hipLaunchKernelGGL(dev::hybrid_merge_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, hits, neural, n, live,
                       any_hit ? 1u : 0u);
    return hipGetLastError();
}

hipError_t query_triangles(const TriArgs& a, const RayBatch& rb, hipStream_t stream)
{
    if (rb.parts)
        hipLaunchKernelGGL(dev::tri_query_kernel<true>, dim3((rb.n + 255) / 256), dim3(256), 0, stream, a, rb);
    else
        hipLaunchKernelGGL(dev::tri_query_kernel<false>, dim3((rb.n + 255) / 256), dim3(256), 0, stream, a, rb);
    return hipGetLastError();
}

}  // namespace nbvh
