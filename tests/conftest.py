import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def nb():
    import nbvh_amd
    return nbvh_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def bench_scene(nb, subdivisions=6, seed=7, n_leaves=256):
    """The benchmark mesh: displaced icosphere, its BVH and N-BVH cut."""
    v, t = nb.icosphere(subdivisions, seed=seed, amplitude=0.1)
    nodes, prims = nb.bvh_build(v, t)
    return v, t, nodes, prims, nb.nbvh_cut(nodes, n_leaves)


@pytest.fixture(scope="session")
def small_scene(nb):
    return bench_scene(nb, subdivisions=4, seed=3, n_leaves=64)


def camera_rays(nb, orc, w, h, eye=(0.0, 0.0, 3.0), sample=0, frame=0):
    cam = nb.camera_look_at(eye, (0, 0, 0), (0, 1, 0), 45.0, w, h)
    return cam, orc.raygen(nb.camera_array(cam), w, h, sample, frame)


def random_rays(n, seed, origin_box=2.5, inside_frac=0.0):
    """Rays from random origins (a fraction inside the unit sphere) to random directions."""
    rng = np.random.default_rng(seed)
    from oracle import RAY_DTYPE
    rays = np.zeros(n, RAY_DTYPE)
    o = rng.uniform(-origin_box, origin_box, (n, 3))
    k = int(n * inside_frac)
    o[:k] = rng.uniform(-0.6, 0.6, (k, 3))
    target = rng.uniform(-0.8, 0.8, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"] = o.astype(np.float32)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = np.float32(1e-4)
    rays["tmax"] = np.float32(3.4028234663852886e38)
    return rays
