"""The FLIP restatement (oracle/flip.py), pinned by what the published method fixes: textbook
CIELab values of the sRGB primaries, zero error for identical images, the closed form for two
uniform images (filtering leaves them unchanged and they have no features), symmetry, locality
(far from a step the error is the uniform closed form), kernel normalisation, and the reference's
viewing geometry."""
import numpy as np
import pytest

from oracle import flip as F


def test_ppd_of_the_reference_monitor():
    assert F.PPD_REFERENCE == pytest.approx(67.0206432, rel=1e-7)
    assert F.csf_radius(F.PPD_REFERENCE) == 10 and F.feature_radius(F.PPD_REFERENCE) == 9


def test_colour_tables_identities():
    """The tabulated constants agree with each other to their printed digits: the illuminant is
    the XYZ of linear white (the matrix's row sums), its reciprocal and the inverse matrix are
    tabulated to 9-10 digits, and Y of white is exactly 1."""
    assert np.allclose(F.M_RGB2XYZ.sum(axis=1), F.WHITE, rtol=0, atol=1e-9)
    assert F.M_RGB2XYZ[1].sum() == 1.0
    assert np.allclose(F.WHITE * F.INV_WHITE, 1.0, rtol=0, atol=3e-8)
    assert np.allclose(F.M_XYZ2RGB @ F.M_RGB2XYZ, np.eye(3), rtol=0, atol=5e-8)
    # sRGB / D65 luminance coefficients (ITU-R BT.709) in the Y row
    assert np.allclose(F.M_RGB2XYZ[1], [0.2126, 0.7152, 0.0722], atol=1e-4)


def test_lab_of_primaries_textbook():
    lab = F.rgb_to_lab(np.eye(3))
    # CIELab (D65) of the sRGB primaries, as tabulated for sRGB (computed with the 7-digit sRGB
    # matrix, which differs from the exact-ratio table in the 5th digit)
    assert np.allclose(lab[0], [53.2408, 80.0925, 67.2032], atol=1e-2)
    assert np.allclose(lab[1], [87.7347, -86.1827, 83.1793], atol=1e-2)
    assert np.allclose(lab[2], [32.2970, 79.1875, -107.8602], atol=1e-2)
    assert np.allclose(F.rgb_to_lab(np.ones(3)), [100, 0, 0], atol=1e-5)
    assert np.allclose(F.rgb_to_lab(np.zeros(3)), [0, 0, 0], atol=1e-9)
    # YyCxCz round trip (to the tables' digits) and sRGB transfer round trip
    x = np.random.default_rng(0).uniform(0, 1, (100, 3))
    assert np.allclose(F.ycxcz_to_rgb(F.rgb_to_ycxcz(x)), x, rtol=0, atol=1e-7)
    assert np.allclose(F.srgb_to_linear(F.linear_to_srgb(x)), x, atol=1e-12)


def test_kernels_normalised():
    for k in F.csf_kernels(F.PPD_REFERENCE):
        assert k.sum() == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(k, k.T) and np.allclose(k, k[::-1, ::-1])
    e, p = F.feature_kernels(F.PPD_REFERENCE)
    for k in (e, p):
        assert k[k > 0].sum() == pytest.approx(1.0) and k[k < 0].sum() == pytest.approx(-1.0)
    assert np.allclose(e, -e[:, ::-1]) and np.allclose(p, p[:, ::-1])


def uniform(h, w, c):
    return np.broadcast_to(np.asarray(c, np.float64), (h, w, 3)).copy()


def closed_form(c1, c2):
    lab = lambda c: F.hunt(F.rgb_to_lab(F.srgb_to_linear(np.clip(F.linear_to_srgb(np.asarray(c, float)), 0, 1))))
    return float(F.remap_color(F.hyab(lab(c1), lab(c2)) ** F.QC))


def test_identical_images_zero():
    img = np.random.default_rng(1).uniform(0, 1, (40, 50, 3))
    assert (F.flip_map(img, img) == 0).all()


def test_uniform_images_closed_form():
    for c1, c2 in [((0, 0, 0), (1, 1, 1)), ((0.2, 0.3, 0.4), (0.25, 0.3, 0.4)), ((0, 1, 0), (0, 0, 1))]:
        m = F.flip_map(uniform(30, 30, c1), uniform(30, 30, c2))
        assert np.allclose(m, closed_form(c1, c2), atol=1e-12)
    # black against white: HyAB = 100 (lightness only)
    cm = F.cmax()
    de = 100 ** F.QC
    assert closed_form((0, 0, 0), (1, 1, 1)) == pytest.approx(F.PT + (de - F.PC * cm) / (cm - F.PC * cm) * (1 - F.PT))
    # green against blue is the normalisation's maximum: remapped to exactly 1
    assert closed_form((0, 1, 0), (0, 0, 1)) == pytest.approx(1.0, abs=1e-12)


def test_symmetry_and_range():
    rng = np.random.default_rng(2)
    a, b = rng.uniform(0, 1, (32, 48, 3)), rng.uniform(0, 1, (32, 48, 3))
    m = F.flip_map(a, b)
    assert np.allclose(m, F.flip_map(b, a), atol=1e-12)
    assert (m >= 0).all() and (m <= 1 + 1e-12).all()


def test_step_is_local():
    h, w = 40, 120
    ref = uniform(h, w, (0.2, 0.2, 0.2))
    test = ref.copy()
    test[:, 60:] = 0.6  # a vertical step in the test image only
    m = F.flip_map(ref, test)
    r = max(F.csf_radius(F.PPD_REFERENCE), F.feature_radius(F.PPD_REFERENCE))
    # left of the step (beyond the kernels' reach): identical images -> 0
    assert (m[:, :60 - r] == 0).all()
    # right of it: uniform 0.2 against uniform 0.6
    assert np.allclose(m[:, 60 + r + 1:], closed_form((0.2,) * 3, (0.6,) * 3), atol=1e-12)
    # at the step the feature term raises the error above both sides' colour error
    assert m[:, 59:61].max() > closed_form((0.2,) * 3, (0.6,) * 3)


def test_accumulation_buffer():
    acc = np.zeros((2, 2, 4))
    acc[0, 0] = [2.0, 4.0, 6.0, 2.0]
    out = F.accum_to_linear(acc)
    assert np.allclose(out[0, 0], [1, 2, 3]) and (out[1, 1] == 0).all()
