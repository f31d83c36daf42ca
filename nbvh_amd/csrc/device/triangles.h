// triangles.h -- closest / any hit over a scene's triangle BVH (the fallback query, the segment
// labels and the training-data sampler share it).
#pragma once

#include "common.h"

namespace nbvh {
namespace dev {

// CudaBVH2::intersect<ShadowRay> (include/cuda/accel/bvh2.cuh:13-88): nearer child first, the
// farther pushed; leaf primitives tested in order, closest hit shrinks tmax; any hit returns.
struct TriResult {
    bool found;
    float t, u, v;
    uint32_t prim, flags;
};

__device__ __forceinline__ TriResult tri_trace(const TriArgs& a, const Ray& ray, bool any_hit)
{
    const TriSetup s = tri_setup(ray);
    float tmax = ray.tmax;
    uint32_t stack[kTriStackSize];
    int ss = 0;
    uint32_t ni = 0;
    TriResult res{false, FLT_MAX, 0.f, 0.f, 0xFFFFFFFFu, 0u};
    while (true) {
        const Box nd = load_node(a.nodes, ni);
        if (nd.prim_count > 0u) {
            for (uint32_t p = nd.left; p < nd.left + nd.prim_count; p++) {
                const uint32_t tri = a.prims[p];
                const uint32_t i0 = a.indices[3 * tri], i1 = a.indices[3 * tri + 1], i2 = a.indices[3 * tri + 2];
                float t, u, v;
                if (tri_intersect(ray, s, a.verts + 3 * (size_t)i0, a.verts + 3 * (size_t)i1, a.verts + 3 * (size_t)i2,
                                  tmax, t, u, v)) {
                    res.found = true;
                    tmax      = t;
                    res.t     = t;
                    res.prim  = tri;
                    res.u     = u;
                    res.v     = v;
                    if (any_hit) return res;
                }
            }
            if (ss == 0) break;
            ni = stack[--ss];
            continue;
        }
        uint32_t c1 = nd.left, c2 = nd.left + 1u;
        float m1 = box_enter(load_node(a.nodes, c1), ray, tmax);
        float m2 = box_enter(load_node(a.nodes, c2), ray, tmax);
        if (m1 > m2) {
            const float tm = m1;
            m1 = m2;
            m2 = tm;
            const uint32_t tc = c1;
            c1 = c2;
            c2 = tc;
        }
        if (m1 == FLT_MAX) {
            if (ss == 0) break;
            ni = stack[--ss];
        } else {
            ni = c1;
            if (m2 != FLT_MAX) {
                if (ss == kTriStackSize) {
                    res.flags |= NBVH_FLAG_STACK_OVERFLOW;
                    break;
                }
                stack[ss++] = c2;
            }
        }
    }
    return res;
}

// Network targets of a segment (generate_single_bvh_network_output_reference,
// include/cuda/neural/neural_bvh_module_common.cuh:783-861, LocalT depth :586-629): closest
// triangle hit of the ray clipped to [lo, hi); visibility 1/0, depth (t - lo) / ((hi - lo) +
// 1e-3) (double literal as in the reference), normal (n + 1) / 2 with n the normalised geometric
// normal (no vertex normals), albedo the default base colour 0.9; all zero without a hit.
// Returns whether the segment holds a triangle.
__device__ __forceinline__ bool segment_targets(const TriArgs& a, Ray ray, float lo, float hi, float out[8])
{
    ray.tmin = lo;
    ray.tmax = hi;
    const TriResult r = tri_trace(a, ray, false);
    for (int j = 0; j < 8; j++) out[j] = 0.f;
    if (!r.found) return false;
    out[0] = 1.f;
    const float seg = hi - lo;
    out[1] = (float)((double)(r.t - lo) / ((double)seg + 1e-3));
    const float* v0 = a.verts + 3 * (size_t)a.indices[3 * (size_t)r.prim];
    const float* v1 = a.verts + 3 * (size_t)a.indices[3 * (size_t)r.prim + 1];
    const float* v2 = a.verts + 3 * (size_t)a.indices[3 * (size_t)r.prim + 2];
    const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    const float nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float inv = 1.f / sqrtf((nr[0] * nr[0] + nr[1] * nr[1]) + nr[2] * nr[2]);
    for (int c = 0; c < 3; c++) {
        out[2 + c] = (nr[c] * inv + 1.f) * 0.5f;
        out[5 + c] = 0.9f;
    }
    return true;
}

}  // namespace dev
}  // namespace nbvh
