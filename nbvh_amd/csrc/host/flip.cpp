// flip.cpp -- nbvh_flip: the LDR-FLIP error of a rendering against a reference rendering
// (kernels/flip.hip), the image metric of the reference's validation mode
// (compute_and_display_flip_error, src/cuda/neural/neural_bvh_renderer.cu:40-92). The host
// computes the method's kernels and colour constants in FP64 (Andersson et al., HPG 2020).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <vector>

#include "../kernels/launch.h"
#include "context.h"

using namespace nbvh;

namespace {

constexpr double kPi = 3.14159265358979323846;

// The colour tables of the method's reference implementation (figures/flip/flip_api.py:155-205):
// linear sRGB (D65) -> XYZ as exact ratios, its tabulated inverse, the D65 illuminant and its
// tabulated reciprocal (XYZ is normalised by multiplying with the reciprocal).
constexpr double kRgb2Xyz[9] = {10135552.0 / 24577794, 8788810.0 / 24577794, 4435075.0 / 24577794,
                                2613072.0 / 12288897,  8788810.0 / 12288897, 887015.0 / 12288897,
                                1425312.0 / 73733382,  8788810.0 / 73733382, 70074185.0 / 73733382};
constexpr double kXyz2Rgb[9] = {3.241003275,  -1.537398934, -0.498615861, -0.969224334, 1.875930071,
                                0.041554224,  0.055639423,  -0.204011202, 1.057148933};
constexpr double kWhite[3] = {0.950428545, 1.0, 1.088900371};
constexpr double kInvWhite[3] = {1.052156925, 1.0, 0.918357670};

double lab_f(double t)
{
    const double d = 6.0 / 29.0;
    return t > d * d * d ? std::cbrt(t) : t / (3 * d * d) + 4.0 / 29.0;
}

// Hunt-adjusted CIELab of a linear RGB colour.
void hunt_lab(const double rgb[3], double lab[3])
{
    double n[3];
    for (int r = 0; r < 3; r++)
        n[r] = (kRgb2Xyz[3 * r] * rgb[0] + kRgb2Xyz[3 * r + 1] * rgb[1] + kRgb2Xyz[3 * r + 2] * rgb[2]) * kInvWhite[r];
    const double L = 116 * lab_f(n[1]) - 16;
    lab[0] = L;
    lab[1] = 0.01 * L * 500 * (lab_f(n[0]) - lab_f(n[1]));
    lab[2] = 0.01 * L * 200 * (lab_f(n[1]) - lab_f(n[2]));
}

}  // namespace

NBVH_API int32_t nbvh_flip(nbvh_ctx* ctx, const float* d_reference, const float* d_test, uint32_t width,
                           uint32_t height, float ppd, float* d_error_map, float* mean_error, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && d_reference && d_test && mean_error, "NULL argument");
        require(width > 0 && height > 0 && (uint64_t)width * height < (1ull << 31), "image size out of range");
        const double p = ppd > 0.f ? (double)ppd : 0.7 * (3840.0 / 0.7) * (kPi / 180.0);
        const int rc = (int)std::ceil(3 * std::sqrt(0.04 / (2 * kPi * kPi)) * p);
        const double sd = 0.5 * 0.082 * p;
        const int rf = (int)std::ceil(3 * sd);
        if (rc > kFlipMaxRadius || rf > kFlipMaxRadius)
            fail(NBVH_ERR_UNSUPPORTED, "pixels per degree too large for the FLIP kernels' radius limit");
        const int nc = 2 * rc + 1, nf = 2 * rf + 1;
        // contrast sensitivity kernels: a1 sqrt(pi/b1) e^(-pi^2 d^2/b1) + a2 sqrt(pi/b2) e^(-pi^2 d^2/b2)
        const double csf[3][4] = {{1.0, 0.0047, 0.0, 1e-5}, {1.0, 0.0053, 0.0, 1e-5}, {34.1, 0.04, 13.5, 0.025}};
        std::vector<float> w((size_t)3 * nc * nc + 2 * (size_t)nf * nf);
        for (int c = 0; c < 3; c++) {
            std::vector<double> k((size_t)nc * nc);
            double sum = 0;
            for (int y = -rc; y <= rc; y++)
                for (int x = -rc; x <= rc; x++) {
                    const double d2 = ((double)x * x + (double)y * y) / (p * p);
                    const double v = csf[c][0] * std::sqrt(kPi / csf[c][1]) * std::exp(-kPi * kPi * d2 / csf[c][1]) +
                                     csf[c][2] * std::sqrt(kPi / csf[c][3]) * std::exp(-kPi * kPi * d2 / csf[c][3]);
                    k[(size_t)(y + rc) * nc + (x + rc)] = v;
                    sum += v;
                }
            for (size_t i = 0; i < k.size(); i++) w[(size_t)c * nc * nc + i] = (float)(k[i] / sum);
        }
        // edge (first derivative) and point (second derivative) detectors along x, each kernel's
        // positive and negative weights normalised to +1 and -1
        for (int kind = 0; kind < 2; kind++) {
            std::vector<double> k((size_t)nf * nf);
            double pos = 0, neg = 0;
            for (int y = -rf; y <= rf; y++)
                for (int x = -rf; x <= rf; x++) {
                    const double g = std::exp(-((double)x * x + (double)y * y) / (2 * sd * sd));
                    const double v = kind == 0 ? -x * g : ((double)x * x / (sd * sd) - 1) * g;
                    k[(size_t)(y + rf) * nf + (x + rf)] = v;
                    if (v > 0) pos += v;
                    if (v < 0) neg -= v;
                }
            float* dst = w.data() + 3 * (size_t)nc * nc + (size_t)kind * nf * nf;
            for (size_t i = 0; i < k.size(); i++) dst[i] = (float)(k[i] > 0 ? k[i] / pos : (k[i] < 0 ? k[i] / neg : 0.0));
        }
        double g[3] = {0, 1, 0}, b[3] = {0, 0, 1}, lg[3], lb[3];
        hunt_lab(g, lg);
        hunt_lab(b, lb);
        const double cmax = std::pow(std::fabs(lg[0] - lb[0]) + std::hypot(lg[1] - lb[1], lg[2] - lb[2]), 0.7);
        const hipStream_t st = (hipStream_t)stream;
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        const size_t n = (size_t)width * height, tiles = flip_tiles(width, height);
        const size_t o_prep = 0, o_part = o_prep + 2 * n * 16, o_mean = o_part + tiles * 8, o_w = o_mean + 256;
        char* mem = nullptr;
        hip_check(hipMallocAsync((void**)&mem, o_w + w.size() * 4, st), "hipMallocAsync");
        double mean = 0;
        try {
            hip_check(hipMemcpyAsync(mem + o_w, w.data(), w.size() * 4, hipMemcpyHostToDevice, st), "weights H2D");
            FlipArgs a{};
            a.ref = d_reference;
            a.test = d_test;
            a.w = width;
            a.h = height;
            a.rc = rc;
            a.rf = rf;
            a.csf = reinterpret_cast<const float*>(mem + o_w);
            a.edge = a.csf + 3 * (size_t)nc * nc;
            a.point = a.edge + (size_t)nf * nf;
            a.cmax = (float)cmax;
            for (int i = 0; i < 9; i++) {
                a.rgb2xyz[i] = (float)kRgb2Xyz[i];
                a.xyz2rgb[i] = (float)kXyz2Rgb[i];
            }
            for (int r = 0; r < 3; r++) {
                a.white[r] = (float)kWhite[r];
                a.inv_white[r] = (float)kInvWhite[r];
            }
            a.prep_ref = reinterpret_cast<float*>(mem + o_prep);
            a.prep_test = a.prep_ref + 4 * n;
            a.map = d_error_map;
            a.partials = reinterpret_cast<double*>(mem + o_part);
            hip_check(flip_error(a, reinterpret_cast<double*>(mem + o_mean), st), "flip launch");
            hip_check(hipMemcpyAsync(&mean, mem + o_mean, 8, hipMemcpyDeviceToHost, st), "mean D2H");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        } catch (...) {
            (void)hipFreeAsync(mem, st);
            throw;
        }
        hip_check(hipFreeAsync(mem, st), "hipFreeAsync");
        *mean_error = (float)mean;
    });
}
