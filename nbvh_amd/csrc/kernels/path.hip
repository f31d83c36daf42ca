// path.hip -- the shading step of the path-trace driver (nbvh_path_trace).
//
// One thread per live path after each depth's closest-hit query. It follows
// hybrid_scatter_and_get_next_tlas_traversal_step (src/cuda/neural/path/path_inference_kernels.cu:
// 224-429) for diffuse materials:
//   no hit               illumination += throughput * background(ray); the path ends;
//   depth == max_depth   the path ends;
//   otherwise            throughput *= albedo (Lambertian: bsdf * |cos| / pdf = albedo); a zero
//                        pdf or albedo ends the path; Russian roulette (rr_terminate,
//                        include/cuda/utils/rendering.cuh:24-40) may end it; the next ray leaves
//                        from hit + offset * normal in a cosine-distributed direction.
// A path that ends writes its pixel (accumulate_pixel_value,
// include/cuda/neural/neural_bvh_renderer_kernels.cuh:81-110: rgb += radiance, w += 1; sample 0
// overwrites). A path that goes on is appended to the next depth's arrays with an atomic; its
// random numbers depend on (sample, depth, pixel) only, so the frame does not depend on the order
// of the compacted arrays.
#include <hip/hip_runtime.h>

#include "../device/sampling.h"
#include "launch.h"

namespace nbvh {
namespace dev {

__device__ __forceinline__ void background_eval(const PathArgs& a, float dy, float out[3])
{
    if (a.background == NBVH_BACKGROUND_SKY) {
        const float t = 0.5f * (dy + 1.f);
        out[0] = (1.f - t) * 1.f + t * 0.5f;
        out[1] = (1.f - t) * 1.f + t * 0.7f;
        out[2] = (1.f - t) * 1.f + t * 1.f;
    } else {
        out[0] = a.background_color[0];
        out[1] = a.background_color[1];
        out[2] = a.background_color[2];
    }
}

__device__ __forceinline__ void splat(const PathArgs& a, uint32_t pixel, const float L[3])
{
    float4* px = reinterpret_cast<float4*>(a.accum) + pixel;
    float4 v = make_float4(L[0], L[1], L[2], 1.f);
    if (a.sample != 0u) {
        const float4 old = *px;
        v = make_float4(old.x + v.x, old.y + v.y, old.z + v.z, old.w + v.w);
    }
    *px = v;
}

__global__ __launch_bounds__(256) void path_init_kernel(PathState* __restrict__ st, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    st[i] = PathState{i, {1.f, 1.f, 1.f}, {0.f, 0.f, 0.f}, 0u};
}

__global__ __launch_bounds__(256) void path_shade_kernel(PathArgs a, uint32_t depth, uint32_t n,
                                                         const uint32_t* __restrict__ live,
                                                         const nbvh_ray* __restrict__ rays,
                                                         const nbvh_hit* __restrict__ hits,
                                                         const PathState* __restrict__ st,
                                                         nbvh_ray* __restrict__ next_rays,
                                                         PathState* __restrict__ next_st,
                                                         uint32_t* __restrict__ counter)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (live && *live < n ? *live : n)) return;
    const Ray r = load_ray(rays, i);
    const nbvh_hit h = hits[i];
    PathState s = st[i];
    const uint32_t kind = h.info & 0xFFu;
    if (kind == NBVH_HIT_NONE) {
        float bg[3];
        background_eval(a, r.dy, bg);
        for (int c = 0; c < 3; c++) s.illum[c] = s.illum[c] + s.thr[c] * bg[c];
        splat(a, s.pixel, s.illum);
        return;
    }
    if (depth >= a.max_depth) {
        splat(a, s.pixel, s.illum);
        return;
    }
    float nrm[3];
    if (!shading_normal(a.tri, h, r.dx, r.dy, r.dz, nrm)) {
        splat(a, s.pixel, s.illum);
        return;
    }
    const float* alb = kind == NBVH_HIT_NEURAL ? h.albedo : a.env_albedo;
    uint32_t rng = lcg_seed(a.sample, depth, s.pixel);
    const float u0 = lcg_float(rng), u1 = lcg_float(rng);
    float wi[3];
    const float cos_i = cosine_direction(u0, u1, nrm, wi);
    for (int c = 0; c < 3; c++) s.thr[c] = s.thr[c] * alb[c];
    if (!(cos_i > 0.f) || (alb[0] == 0.f && alb[1] == 0.f && alb[2] == 0.f)) {
        splat(a, s.pixel, s.illum);
        return;
    }
    if (a.rr_start_depth != 0u && depth >= a.rr_start_depth) {
        const float survive = fminf(1.f, fmaxf(fmaxf(s.thr[0], s.thr[1]), s.thr[2]));
        if (lcg_float(rng) < survive) {
            for (int c = 0; c < 3; c++) s.thr[c] = s.thr[c] / survive;
        } else {
            splat(a, s.pixel, s.illum);
            return;
        }
    }
    float p[3] = {r.ox + h.t * r.dx, r.oy + h.t * r.dy, r.oz + h.t * r.dz};
    for (int c = 0; c < 3; c++) p[c] = p[c] + nrm[c] * a.normal_offset;
    const uint32_t j = atomicAdd(counter, 1u);
    float4* o = reinterpret_cast<float4*>(next_rays + j);
    o[0] = make_float4(p[0], p[1], p[2], 1e-4f);
    o[1] = make_float4(wi[0], wi[1], wi[2], FLT_MAX);
    next_st[j] = s;
}

}  // namespace dev

hipError_t path_init(PathState* st, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::path_init_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, st, n);
    return hipGetLastError();
}

hipError_t path_shade(const PathArgs& a, uint32_t depth, uint32_t n, const uint32_t* live, const nbvh_ray* rays,
                      const nbvh_hit* hits, const PathState* st, nbvh_ray* next_rays, PathState* next_st,
                      uint32_t* counter, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::path_shade_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, depth, n, live, rays,
                       hits, st, next_rays, next_st, counter);
    return hipGetLastError();
}

}  // namespace nbvh
