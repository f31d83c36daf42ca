// flip.hip -- LDR-FLIP image error of a rendering against a reference rendering, the metric of the
// reference's validation mode (compute_and_display_flip_error,
// src/cuda/neural/neural_bvh_renderer.cu:40-92), written from the published method (Andersson et
// al., "FLIP: A Difference Evaluator for Alternating Images", HPG 2020) with the steps and
// constants of the method's reference implementation (figures/flip/flip_api.py:439-497):
//
//   flip_prepare_kernel  per pixel of both accumulation buffers (rgb sum, sample count): mean
//                        linear RGB -> sRGB clamped to [0, 1] -> linear RGB -> YyCxCz and Y / Yn;
//   flip_error_kernel    16 x 16 pixel tiles with a halo of the larger kernel radius in LDS
//                        (edge pixels replicated): the contrast-sensitivity filtered YyCxCz of both
//                        images -> linear RGB clamped -> CIELab -> Hunt -> HyAB^0.7 remapped (colour
//                        error); edge and point responses of Y / Yn (feature error); the pixel's
//                        error dEc^(1 - dEf); per-tile sums of the errors;
//   flip_mean_kernel     the tile sums added in tile order (one thread), / pixel count.
// The kernels' weights are computed on the host (FP64, rounded to FP32). FP32 arithmetic.
#include <hip/hip_runtime.h>

#include "../nbvh_internal.h"
#include "launch.h"

namespace nbvh {
namespace dev {

constexpr int kFlipTile = 16;

__device__ __forceinline__ float srgb_encode(float c)
{
    return c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.f / 2.4f) - 0.055f;
}
__device__ __forceinline__ float srgb_decode(float c)
{
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// XYZ of linear sRGB (D65 primaries) normalised by the illuminant (times its reciprocal).
__device__ __forceinline__ void rgb_to_xyz_n(const FlipArgs& a, const float rgb[3], float xyz[3])
{
    for (int r = 0; r < 3; r++)
        xyz[r] = (a.rgb2xyz[3 * r] * rgb[0] + a.rgb2xyz[3 * r + 1] * rgb[1] + a.rgb2xyz[3 * r + 2] * rgb[2]) * a.inv_white[r];
}

__device__ __forceinline__ float lab_f(float t)
{
    const float d = 6.f / 29.f;
    return t > d * d * d ? cbrtf(t) : t / (3.f * d * d) + 4.f / 29.f;
}

// Hunt-adjusted CIELab of a YyCxCz colour after conversion to linear RGB clamped to [0, 1].
__device__ __forceinline__ void ycc_to_hunt_lab(const FlipArgs& a, const float ycc[3], float lab[3])
{
    const float y = (ycc[0] + 16.f) / 116.f;
    const float xyz[3] = {(ycc[1] / 500.f + y) * a.white[0], y * a.white[1], (y - ycc[2] / 200.f) * a.white[2]};
    float rgb[3];
    for (int r = 0; r < 3; r++) {
        const float v = a.xyz2rgb[3 * r] * xyz[0] + a.xyz2rgb[3 * r + 1] * xyz[1] + a.xyz2rgb[3 * r + 2] * xyz[2];
        rgb[r] = fminf(fmaxf(v, 0.f), 1.f);
    }
    float n[3];
    rgb_to_xyz_n(a, rgb, n);
    const float fx = lab_f(n[0]), fy = lab_f(n[1]), fz = lab_f(n[2]);
    const float L = 116.f * fy - 16.f;
    lab[0] = L;
    lab[1] = 0.01f * L * (500.f * (fx - fy));
    lab[2] = 0.01f * L * (200.f * (fy - fz));
}

__global__ __launch_bounds__(256) void flip_prepare_kernel(FlipArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.w * a.h) return;
    for (int img = 0; img < 2; img++) {
        const float4 acc = reinterpret_cast<const float4*>(img == 0 ? a.ref : a.test)[i];
        float rgb[3] = {0.f, 0.f, 0.f};
        if (acc.w > 0.f) {
            rgb[0] = acc.x / acc.w;
            rgb[1] = acc.y / acc.w;
            rgb[2] = acc.z / acc.w;
        }
        for (int c = 0; c < 3; c++) rgb[c] = srgb_decode(fminf(fmaxf(srgb_encode(rgb[c]), 0.f), 1.f));
        float n[3];
        rgb_to_xyz_n(a, rgb, n);
        reinterpret_cast<float4*>(img == 0 ? a.prep_ref : a.prep_test)[i] =
            make_float4(116.f * n[1] - 16.f, 500.f * (n[0] - n[1]), 200.f * (n[1] - n[2]), n[1]);
    }
}

__global__ __launch_bounds__(256) void flip_error_kernel(FlipArgs a)
{
    constexpr int R = kFlipMaxRadius;
    constexpr int S = kFlipTile + 2 * R;
    __shared__ float4 tile[2][S * S];
    __shared__ double sums[256];
    const int tx = threadIdx.x % kFlipTile, ty = threadIdx.x / kFlipTile;
    const int x0 = blockIdx.x * kFlipTile, y0 = blockIdx.y * kFlipTile;
    const int W = (int)a.w, H = (int)a.h;
    for (int k = threadIdx.x; k < S * S; k += blockDim.x) {
        const int gx = min(max(x0 + k % S - R, 0), W - 1), gy = min(max(y0 + k / S - R, 0), H - 1);
        tile[0][k] = reinterpret_cast<const float4*>(a.prep_ref)[(size_t)gy * W + gx];
        tile[1][k] = reinterpret_cast<const float4*>(a.prep_test)[(size_t)gy * W + gx];
    }
    __syncthreads();
    const int px = x0 + tx, py = y0 + ty;
    float err = 0.f;
    if (px < W && py < H) {
        const int rc = a.rc, rf = a.rf, nc = 2 * rc + 1, nf = 2 * rf + 1;
        float lab[2][3], edge[2], point[2];
        for (int img = 0; img < 2; img++) {
            const float4* t = tile[img];
            float f[3] = {0.f, 0.f, 0.f};
            for (int dy = -rc; dy <= rc; dy++) {
                const float4* row = t + (ty + R + dy) * S + (tx + R);
                const float* w0 = a.csf + (dy + rc) * nc + rc;
                for (int dx = -rc; dx <= rc; dx++) {
                    const float4 v = row[dx];
                    f[0] = fmaf(w0[dx], v.x, f[0]);
                    f[1] = fmaf(w0[dx + nc * nc], v.y, f[1]);
                    f[2] = fmaf(w0[dx + 2 * nc * nc], v.z, f[2]);
                }
            }
            ycc_to_hunt_lab(a, f, lab[img]);
            // The detectors' weights sum to zero, so they are applied to Y - Y(centre): the same
            // responses, without the rounding residue of the weights' sum times Y, which the
            // square root of the feature error would magnify on flat regions.
            const float yc = t[(ty + R) * S + (tx + R)].w;
            float ex = 0.f, ey = 0.f, qx = 0.f, qy = 0.f;
            for (int dy = -rf; dy <= rf; dy++) {
                const float4* row = t + (ty + R + dy) * S + (tx + R);
                for (int dx = -rf; dx <= rf; dx++) {
                    const float yv = row[dx].w - yc;
                    const int kx = (dy + rf) * nf + (dx + rf), ky = (dx + rf) * nf + (dy + rf);
                    ex = fmaf(a.edge[kx], yv, ex);
                    ey = fmaf(a.edge[ky], yv, ey);
                    qx = fmaf(a.point[kx], yv, qx);
                    qy = fmaf(a.point[ky], yv, qy);
                }
            }
            edge[img]  = sqrtf(ex * ex + ey * ey);
            point[img] = sqrtf(qx * qx + qy * qy);
        }
        const float dl = fabsf(lab[0][0] - lab[1][0]);
        const float da = lab[0][1] - lab[1][1], db = lab[0][2] - lab[1][2];
        float dec = powf(dl + sqrtf(da * da + db * db), 0.7f);
        const float pcm = 0.4f * a.cmax;
        dec = dec < pcm ? (0.95f / pcm) * dec : 0.95f + (dec - pcm) / (a.cmax - pcm) * 0.05f;
        const float fd = fmaxf(fabsf(edge[0] - edge[1]), fabsf(point[0] - point[1])) * 0.70710678118654752f;
        const float def = sqrtf(fd);
        err = powf(dec, 1.f - def);
        if (a.map) a.map[(size_t)py * W + px] = err;
    }
    sums[threadIdx.x] = (double)err;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sums[threadIdx.x] += sums[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partials[blockIdx.y * gridDim.x + blockIdx.x] = sums[0];
}

__global__ void flip_mean_kernel(const double* __restrict__ partials, uint32_t n, double pixels, double* __restrict__ out)
{
    double s = 0.0;
    for (uint32_t i = 0; i < n; i++) s += partials[i];
    *out = s / pixels;
}

}  // namespace dev

uint32_t flip_tiles(uint32_t w, uint32_t h)
{
    return ((w + dev::kFlipTile - 1) / dev::kFlipTile) * ((h + dev::kFlipTile - 1) / dev::kFlipTile);
}

hipError_t flip_error(const FlipArgs& a, double* mean, hipStream_t stream)
{
    const uint32_t n = a.w * a.h;
    hipLaunchKernelGGL(dev::flip_prepare_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid((a.w + dev::kFlipTile - 1) / dev::kFlipTile, (a.h + dev::kFlipTile - 1) / dev::kFlipTile);
    hipLaunchKernelGGL(dev::flip_error_kernel, grid, dim3(256), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::flip_mean_kernel, dim3(1), dim3(1), 0, stream, a.partials, grid.x * grid.y, (double)n,
                       mean);
    return hipGetLastError();
}

}  // namespace nbvh
