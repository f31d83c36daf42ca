// sampling.h -- random numbers and diffuse sampling shared by the bounce and path kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace nbvh {
namespace dev {

// murmur_mix, murmur_fmix, lcg_seed and lcg_float (get_lcg_rng, lcg_randomf) are in common.h.

// ortho_basis (include/cuda/utils/common.cuh:21-35).
__device__ __forceinline__ void ortho_basis(const float n[3], float t[3], float b[3])
{
    float bt[3] = {0.f, 0.f, 0.f};
    if (n[0] < 0.6f && n[0] > -0.6f)
        bt[0] = 1.f;
    else if (n[1] < 0.6f && n[1] > -0.6f)
        bt[1] = 1.f;
    else if (n[2] < 0.6f && n[2] > -0.6f)
        bt[2] = 1.f;
    else
        bt[0] = 1.f;
    t[0] = bt[1] * n[2] - bt[2] * n[1];
    t[1] = bt[2] * n[0] - bt[0] * n[2];
    t[2] = bt[0] * n[1] - bt[1] * n[0];
    float inv = 1.f / sqrtf((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]);
    for (int c = 0; c < 3; c++) t[c] *= inv;
    b[0] = n[1] * t[2] - n[2] * t[1];
    b[1] = n[2] * t[0] - n[0] * t[2];
    b[2] = n[0] * t[1] - n[1] * t[0];
    inv = 1.f / sqrtf((b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]);
    for (int c = 0; c < 3; c++) b[c] *= inv;
}

// Shading normal of a hit, turned toward the incoming direction d (shading_frame,
// include/cuda/utils/rendering.cuh:11-22): the predicted normal of a neural hit or the geometric
// normal of a triangle hit (triangles of `tri`). Returns false when the hit has no normal source
// (no hit, or a triangle hit without triangle data).
__device__ __forceinline__ bool shading_normal(const TriArgs& tri, const nbvh_hit& h, float dx, float dy, float dz,
                                               float nrm[3])
{
    const uint32_t kind = h.info & 0xFFu;
    if (kind == NBVH_HIT_NEURAL) {
        nrm[0] = h.normal_or_uv[0];
        nrm[1] = h.normal_or_uv[1];
        nrm[2] = h.normal_or_uv[2];
    } else if (kind == NBVH_HIT_TRIANGLE && tri.verts != nullptr) {
        const float* a = tri.verts + 3 * (size_t)tri.indices[3 * (size_t)h.prim_id];
        const float* b = tri.verts + 3 * (size_t)tri.indices[3 * (size_t)h.prim_id + 1];
        const float* c = tri.verts + 3 * (size_t)tri.indices[3 * (size_t)h.prim_id + 2];
        const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
    } else {
        return false;
    }
    const float len = sqrtf((nrm[0] * nrm[0] + nrm[1] * nrm[1]) + nrm[2] * nrm[2]);
    const float sgn = ((nrm[0] * dx + nrm[1] * dy) + nrm[2] * dz) > 0.f ? -1.f : 1.f;
    const float inv = len > 0.f ? sgn / len : 0.f;
    for (int c = 0; c < 3; c++) nrm[c] *= inv;
    if (!(len > 0.f)) nrm[2] = 1.f;
    return true;
}

// Cosine-distributed direction about nrm from two uniforms (sample_lambertian_dir +
// cos_sample_hemisphere: concentric disk mapping, z = sqrt(1 - r^2); material.cuh:101-120,
// 193-201). Returns the local z (cos theta; 0 at the disk's rim, where the pdf vanishes).
__device__ __forceinline__ float cosine_direction(float u0, float u1, const float nrm[3], float wi[3])
{
    const float sx = 2.f * u0 - 1.f, sy = 2.f * u1 - 1.f;
    float dx = 0.f, dy = 0.f;
    if (sx != 0.f || sy != 0.f) {
        float rad, th;
        if (fabsf(sx) > fabsf(sy)) {
            rad = sx;
            th  = 0.785398163397448310f * (sy / sx);
        } else {
            rad = sy;
            th  = 1.570796326794896619f - 0.785398163397448310f * (sx / sy);
        }
        dx = rad * cosf(th);
        dy = rad * sinf(th);
    }
    float hz = sqrtf(fmaxf(0.f, 1.f - dx * dx - dy * dy));
    const float hinv = 1.f / sqrtf((dx * dx + dy * dy) + hz * hz);
    const float hx = dx * hinv, hy = dy * hinv;
    hz *= hinv;
    float t[3], b[3];
    ortho_basis(nrm, t, b);
    for (int c = 0; c < 3; c++) wi[c] = (hx * t[c] + hy * b[c]) + hz * nrm[c];
    return hz;
}

}  // namespace dev
}  // namespace nbvh
