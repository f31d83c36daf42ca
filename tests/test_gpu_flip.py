"""nbvh_flip (LDR-FLIP, FP32 on the GPU) against the FP64 restatement (oracle/flip.py).

Bars: identical images give exactly 0; two uniform images give the method's closed form within
1e-5; on structured images of sizes that are not multiples of the 16-pixel tile, 99.9 % of the
pixels are within 2e-4 of the FP64 error map and every pixel within 2e-2 (FP32 filtering next to
the colour remap's knee and the feature maximum), the mean within 1e-5; and a neural path-traced
frame against the triangle rendering of the same object, as the reference's validation compares
them.
"""
import numpy as np
import pytest

from oracle import flip as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


def accum(rgb, count):
    h, w = rgb.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    out[..., :3] = rgb * count
    out[..., 3] = count
    return out


def structured(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w] / np.array([h, w])[:, None, None]
    img = np.stack([0.6 * x, 0.5 * y, 0.3 + 0.2 * np.sin(9 * x * y)], axis=-1)
    img[h // 3: h // 2, w // 4: w // 2] = (0.9, 0.2, 0.1)  # a block with edges and corners
    img += rng.normal(0, 0.03, img.shape)
    return np.clip(img, 0, 1.5)


def check_against_oracle(nb, ctx, ref_acc, test_acc):
    mean, m = nb.flip(ctx, ref_acc, test_acc)
    exp = F.flip_map(F.accum_to_linear(ref_acc), F.accum_to_linear(test_acc))
    d = np.abs(m.astype(np.float64) - exp)
    assert np.quantile(d, 0.999) < 2e-4 and d.max() < 2e-2, (np.quantile(d, 0.999), d.max())
    assert abs(mean - exp.mean()) < 1e-5, (mean, exp.mean())
    return mean, exp


def test_identical_and_uniform(nb, ctx):
    img = accum(structured(70, 90, 1), 3.0)
    mean, m = nb.flip(ctx, img, img)
    assert mean == 0.0 and (m == 0).all()
    a = accum(np.full((40, 40, 3), 0.2), 1.0)
    b = accum(np.full((40, 40, 3), 0.6), 1.0)
    mean, m = nb.flip(ctx, a, b)
    lab = lambda c: F.hunt(F.rgb_to_lab(np.full(3, c)))  # noqa: E731
    closed = float(F.remap_color(F.hyab(lab(0.2), lab(0.6)) ** F.QC))
    assert np.allclose(m, closed, atol=1e-5) and abs(mean - closed) < 1e-5


@pytest.mark.parametrize("shape", [(77, 123), (64, 64), (17, 300)])
def test_structured_images(nb, ctx, shape):
    h, w = shape
    ref = structured(h, w, 2)
    test = ref + np.random.default_rng(3).normal(0, 0.05, ref.shape)
    test[h // 2:, : w // 3] *= 0.7
    check_against_oracle(nb, ctx, accum(ref, 4.0), accum(np.clip(test, 0, 2), 4.0))


def test_neural_rendering_against_triangle_rendering(nb, ctx):
    v, t = nb.icosphere(5, seed=7, amplitude=0.1)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(nb.network_config(), seed=1, table_range=1e-4)
    mod = nb.NeuralBVHModule(ctx, scene)
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, 128, 96)
    neural, _ = mod.path_trace(cam, 128, 96, samples=2, precision=nb.PREC_FP16, max_depth=3)
    tri, _ = mod.path_trace(cam, 128, 96, samples=2, triangles=True, max_depth=3, environment_albedo=(0.9,) * 3)
    mean, _ = check_against_oracle(nb, ctx, tri, neural)
    assert 0.0 < mean < 1.0
    scene.close()
