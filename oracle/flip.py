"""CPU restatement of LDR-FLIP (test infrastructure only; never used by the product).

The reference compares each neural rendering with the triangle rendering by FLIP
(compute_and_display_flip_error, src/cuda/neural/neural_bvh_renderer.cu:40-92: both accumulation
buffers divided by the sample count, linear RGB converted to sRGB, FLIP at the pixels per degree
of a 0.7 m wide 3840-pixel monitor seen from 0.7 m, mean of the error map). Its FLIP is NVIDIA's
flip-cuda (an empty submodule here) and the Python tool under figures/flip needs OpenCV and
OpenEXR, which are not importable, so this follows the published method (Andersson, Nilsson,
Akenine-Moeller, Oskarsson, Astrom, Fairchild: "FLIP: A Difference Evaluator for Alternating
Images", HPG 2020) with every constant and step as the reference's Python tool has it
(figures/flip/flip_api.py:134-205 colour tables, :275-323 contrast sensitivity kernels, :326-348
filtering and clamp, :350-398 Hunt, HyAB, remap, :400-437 feature kernels, :439-497 LDR-FLIP),
in FP64 where that tool computes in FP32:

  colour  sRGB -> linear RGB -> XYZ (sRGB primaries, D65; the tables of the method's own
          implementation, figures/flip/flip_api.py:134-205) -> YyCxCz; each channel filtered with
          its contrast sensitivity kernel, the sum of Gaussians a1 sqrt(pi/b1) exp(-pi^2 d^2/b1) +
          a2 sqrt(pi/b2) exp(-pi^2 d^2/b2) of the distance d in degrees (achromatic (1, 0.0047,
          0, 1e-5), red-green (1, 0.0053, 0, 1e-5), blue-yellow (34.1, 0.04, 13.5, 0.025)),
          normalised to sum 1, radius ceil(3 sqrt(0.04 / (2 pi^2)) ppd) pixels; back to linear
          RGB clamped to [0, 1]; CIELab; Hunt adjustment (L, 0.01 L a, 0.01 L b); HyAB
          distance |dL| + ||(da, db)||, to the power 0.7; remapped with cmax (the same of green
          against blue), pc 0.4, pt 0.95;
  feature Y / Yn of the unfiltered images; first and second derivatives of a Gaussian of
          standard deviation 0.5 * 0.082 * ppd pixels (radius ceil(3 sd)), each kernel's positive
          and negative weights normalised to sum +1 and -1; x and y responses; edge and point
          magnitudes; dEf = (max(|edge difference|, |point difference|) / sqrt(2))^0.5;
  error   dEc^(1 - dEf) per pixel; the mean over pixels.

Borders replicate the edge pixel. Inputs: linear RGB, clamped to [0, 1] after conversion to sRGB
(LDR-FLIP's input range). The reference's FLIP code is not runnable here (flip-cuda absent, the
Python tool needs OpenCV), so no output of it pins this; the oracle is pinned by the method's
closed forms and the colour tables' own identities (tests/test_oracle_flip.py).
"""
import numpy as np
from scipy import ndimage

PPD_REFERENCE = 0.7 * (3840.0 / 0.7) * (np.pi / 180.0)  # calculatePPD(0.7 m, 3840 px, 0.7 m)

CSF = {"A": (1.0, 0.0047, 0.0, 1e-5), "RG": (1.0, 0.0053, 0.0, 1e-5), "BY": (34.1, 0.04, 13.5, 0.025)}
QC, QF, PC, PT, GW = 0.7, 0.5, 0.4, 0.95, 0.082

# The colour tables of the method's reference implementation (figures/flip/flip_api.py:155-205,
# color_space_transform): linear sRGB -> XYZ as exact ratios, its tabulated inverse, the D65
# illuminant and its tabulated reciprocal (XYZ is normalised by multiplying with the reciprocal).
M_RGB2XYZ = np.array([[10135552 / 24577794, 8788810 / 24577794, 4435075 / 24577794],
                      [2613072 / 12288897, 8788810 / 12288897, 887015 / 12288897],
                      [1425312 / 73733382, 8788810 / 73733382, 70074185 / 73733382]])
M_XYZ2RGB = np.array([[3.241003275, -1.537398934, -0.498615861],
                      [-0.969224334, 1.875930071, 0.041554224],
                      [0.055639423, -0.204011202, 1.057148933]])
WHITE = np.array([0.950428545, 1.0, 1.088900371])
INV_WHITE = np.array([1.052156925, 1.0, 0.918357670])


def linear_to_srgb(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0031308), 1 / 2.4) - 0.055)


def srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, np.power((np.maximum(c, 0.04045) + 0.055) / 1.055, 2.4))


def rgb_to_ycxcz(rgb):
    xyz = (rgb @ M_RGB2XYZ.T) * INV_WHITE
    return np.stack([116 * xyz[..., 1] - 16, 500 * (xyz[..., 0] - xyz[..., 1]), 200 * (xyz[..., 1] - xyz[..., 2])],
                    axis=-1)


def ycxcz_to_rgb(ycc):
    y = (ycc[..., 0] + 16) / 116
    x = ycc[..., 1] / 500 + y
    z = y - ycc[..., 2] / 200
    return (np.stack([x, y, z], axis=-1) * WHITE) @ M_XYZ2RGB.T


def _lab_f(t):
    d = 6.0 / 29.0
    return np.where(t > d ** 3, np.cbrt(t), t / (3 * d * d) + 4.0 / 29.0)


def rgb_to_lab(rgb):
    xyz = (rgb @ M_RGB2XYZ.T) * INV_WHITE
    fx, fy, fz = _lab_f(xyz[..., 0]), _lab_f(xyz[..., 1]), _lab_f(xyz[..., 2])
    return np.stack([116 * fy - 16, 500 * (fx - fy), 200 * (fy - fz)], axis=-1)


def hunt(lab):
    return np.stack([lab[..., 0], 0.01 * lab[..., 0] * lab[..., 1], 0.01 * lab[..., 0] * lab[..., 2]], axis=-1)


def hyab(a, b):
    d = a - b
    return np.abs(d[..., 0]) + np.sqrt(d[..., 1] ** 2 + d[..., 2] ** 2)


def cmax():
    g, b = hunt(rgb_to_lab(np.array([0.0, 1.0, 0.0]))), hunt(rgb_to_lab(np.array([0.0, 0.0, 1.0])))
    return float(hyab(g, b) ** QC)


def remap_color(de):
    cm = cmax()
    return np.where(de < PC * cm, (PT / (PC * cm)) * de, PT + (de - PC * cm) / (cm - PC * cm) * (1 - PT))


def csf_radius(ppd):
    return int(np.ceil(3 * np.sqrt(0.04 / (2 * np.pi ** 2)) * ppd))


def csf_kernels(ppd):
    """The three normalised contrast sensitivity kernels, (2R+1)^2 each, row = y offset."""
    r = csf_radius(ppd)
    y, x = np.mgrid[-r:r + 1, -r:r + 1].astype(np.float64)
    d2 = (x * x + y * y) / (ppd * ppd)
    out = []
    for name in ("A", "RG", "BY"):
        a1, b1, a2, b2 = CSF[name]
        k = a1 * np.sqrt(np.pi / b1) * np.exp(-np.pi ** 2 * d2 / b1) + a2 * np.sqrt(np.pi / b2) * np.exp(-np.pi ** 2 * d2 / b2)
        out.append(k / k.sum())
    return out


def feature_radius(ppd):
    return int(np.ceil(3 * 0.5 * GW * ppd))


def feature_kernels(ppd):
    """(edge_x, point_x) kernels, (2R+1)^2, positive and negative parts normalised to +1 / -1;
    the y kernels are their transposes."""
    sd = 0.5 * GW * ppd
    r = feature_radius(ppd)
    y, x = np.mgrid[-r:r + 1, -r:r + 1].astype(np.float64)
    g = np.exp(-(x * x + y * y) / (2 * sd * sd))
    out = []
    for k in (-x * g, (x * x / (sd * sd) - 1) * g):
        k = k.copy()
        pos, neg = k[k > 0].sum(), -k[k < 0].sum()
        k[k > 0] /= pos
        k[k < 0] /= neg
        out.append(k)
    return out


def _conv(img, k):
    return ndimage.correlate(img, k, mode="nearest")


def flip_map(ref_linear, test_linear, ppd=PPD_REFERENCE):
    """Per-pixel FLIP error of two linear-RGB images [h, w, 3]."""
    ref = srgb_to_linear(np.clip(linear_to_srgb(ref_linear), 0.0, 1.0))
    test = srgb_to_linear(np.clip(linear_to_srgb(test_linear), 0.0, 1.0))
    yr, yt = rgb_to_ycxcz(ref), rgb_to_ycxcz(test)
    ks = csf_kernels(ppd)
    fr = np.stack([_conv(yr[..., c], ks[c]) for c in range(3)], axis=-1)
    ft = np.stack([_conv(yt[..., c], ks[c]) for c in range(3)], axis=-1)
    lr = hunt(rgb_to_lab(np.clip(ycxcz_to_rgb(fr), 0.0, 1.0)))
    lt = hunt(rgb_to_lab(np.clip(ycxcz_to_rgb(ft), 0.0, 1.0)))
    dec = remap_color(hyab(lr, lt) ** QC)
    ek, pk = feature_kernels(ppd)

    def features(ycc):
        yn = (ycc[..., 0] + 16) / 116
        e = np.hypot(_conv(yn, ek), _conv(yn, ek.T))
        p = np.hypot(_conv(yn, pk), _conv(yn, pk.T))
        return e, p

    er, pr = features(yr)
    et, pt = features(yt)
    def_ = (np.maximum(np.abs(er - et), np.abs(pr - pt)) / np.sqrt(2)) ** QF
    return np.power(dec, 1 - def_)


def flip_mean(ref_linear, test_linear, ppd=PPD_REFERENCE):
    return float(flip_map(ref_linear, test_linear, ppd).mean())


def accum_to_linear(accum):
    """Accumulation buffer [h, w, 4] (rgb sum, sample count) -> mean linear RGB."""
    a = np.asarray(accum, np.float64)
    n = np.maximum(a[..., 3:4], 1.0)
    return np.where(a[..., 3:4] > 0, a[..., :3] / n, 0.0)
