"""Decode options of the oracle's neural query (WorldT depth, linear depth activation, depth
interpolation; include/cuda/neural/neural_bvh_module_common.cuh:23-84 and
src/cuda/neural/path/path_inference_kernels.cu:89-121), pinned by closed forms.

A network whose tables hold 1 everywhere and whose MLP passes one hidden unit (the first feature,
1 up to rounding) through to the outputs gives the same logits c_j for every segment, so the decoded
depths follow from the geometry alone: WorldT's distance is u * (span + 1e-3) + span_min on the
ray's span through the scene box, whatever leaf the segment came from; the linear activation maps
sigmoid(c) to clamp((sigmoid(c) - 0.45) * 10, 0, 1); with interpolation the closest-hit distance
is the fold of the evaluated segments' LocalT depths (averaged when within the epsilon), and equal
normals stay that normal.
"""
import numpy as np
import pytest

from conftest import bench_scene

import oracle


def constant_network(nb, orc, cfg, logits, bmin, bmax):
    """Tables of ones; MLP with hidden unit 0 = feature 0 (= 1 up to rounding) carried through the
    hidden layers; output j = logits[j] * unit 0."""
    probe = orc.Net(cfg, np.zeros(1, np.float32), np.zeros(1, np.float32), bmin, bmax)
    n_params = int(probe.levels()["n_params"])
    tables = np.ones(n_params * cfg.grid.n_features_per_level, np.float32)
    K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level
    H = cfg.n_neurons
    mats = []
    w1 = np.zeros((H, K0), np.float32)
    w1[0, 0] = 1.0
    mats.append(w1)
    for _ in range(cfg.n_hidden_layers - 1):
        w = np.zeros((H, H), np.float32)
        w[0, 0] = 1.0
        mats.append(w)
    wo = np.zeros((cfg.n_outputs, H), np.float32)
    wo[:, 0] = logits
    mats.append(wo)
    mlp = np.concatenate([m.reshape(-1) for m in mats])
    return orc.Net(cfg, tables, mlp, bmin, bmax)


def sig(x):
    return 1.0 / (1.0 + np.exp(-x))


@pytest.fixture(scope="module")
def scene(nb):
    v, t, nodes, prims, nbvh = bench_scene(nb, subdivisions=4, seed=3, n_leaves=64)
    bmin, bmax = v.min(axis=0), v.max(axis=0)
    rng = np.random.default_rng(8)
    rays = np.zeros(400, nb.RAY_DTYPE)
    o = rng.normal(size=(400, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 2.5
    target = rng.uniform(-0.3, 0.3, (400, 3))
    d = target - o
    rays["o"] = o.astype(np.float32)
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, np.float32(3.4e38)
    return dict(nbvh=nbvh, bmin=bmin, bmax=bmax, rays=rays)


LOGITS = np.array([5.0, 0.3, 0.8, -0.4, 1.1, 0.2, -0.7, 0.5], np.float32)


def span(bmin, bmax, rays):
    out = np.zeros((len(rays), 2), np.float32)
    for i, r in enumerate(rays):
        out[i] = oracle.slab(bmin, bmax, r)[1:]
    return out


def test_world_t_depth_closed_form(nb, scene):
    cfg = nb.network_config(depth_mode=nb.DEPTH_WORLD_T)
    net = constant_network(nb, oracle, cfg, LOGITS, scene["bmin"], scene["bmax"])
    hits = net.neural_query(scene["nbvh"], scene["rays"], mode=oracle.FP64)
    hit = nb.hit_kind(hits) == 1
    assert hit.mean() > 0.5
    sp = span(scene["bmin"], scene["bmax"], scene["rays"][hit])
    u = sig(float(LOGITS[1]))
    expected = (u * ((sp[:, 1] - sp[:, 0]).astype(np.float64) + 1e-3) + sp[:, 0]).astype(np.float32)
    assert np.array_equal(hits["t"][hit], expected)
    # LocalT with the same network: the first segment's own span instead (differs)
    local = constant_network(nb, oracle, nb.network_config(), LOGITS, scene["bmin"], scene["bmax"])
    lh = local.neural_query(scene["nbvh"], scene["rays"], mode=oracle.FP64)
    assert np.array_equal(nb.hit_kind(lh) == 1, hit) and not np.array_equal(lh["t"][hit], hits["t"][hit])


def test_linear_depth_activation_closed_form(nb, scene):
    cfg = nb.network_config(depth_mode=nb.DEPTH_WORLD_T, depth_linear_activation=True)
    lin = lambda c: float(np.clip((sig(float(np.float32(c))) - float(np.float32(0.45))) * 10, 0, 1))  # noqa: E731
    for c1, u in ((0.3, lin(0.3)), (0.5, lin(0.5)), (-2.0, 0.0), (3.0, 1.0)):
        logits = LOGITS.copy()
        logits[1] = c1
        net = constant_network(nb, oracle, cfg, logits, scene["bmin"], scene["bmax"])
        hits = net.neural_query(scene["nbvh"], scene["rays"], mode=oracle.FP64)
        hit = nb.hit_kind(hits) == 1
        sp = span(scene["bmin"], scene["bmax"], scene["rays"][hit])
        expected = (u * ((sp[:, 1] - sp[:, 0]).astype(np.float64) + 1e-3) + sp[:, 0]).astype(np.float32)
        assert np.array_equal(hits["t"][hit], expected), c1


@pytest.mark.parametrize("mode", [oracle.FP32, oracle.FP64])
def test_depth_interpolation_fold(nb, scene, mode):
    """eps large enough that every later segment is within it: the closest-hit distance is the fold
    t <- (t + d) / 2 over the evaluated segments' LocalT depths (FP64 mode exactly; FP32 mode, whose
    network arithmetic makes the unit output 1 only up to rounding, within 1e-6); without
    interpolation the minimum."""
    cfg = nb.network_config(interpolate_depth=True, depth_interpolation_eps=10.0)
    net = constant_network(nb, oracle, cfg, LOGITS, scene["bmin"], scene["bmax"])
    hits, log = net.neural_query(scene["nbvh"], scene["rays"], mode=mode, log_segments=True)
    u = sig(np.float64(LOGITS[1]))
    folded = 0
    for i in range(len(scene["rays"])):
        segs = log[log["ray"] == i]
        t = None
        for s in segs:
            mint, maxt = np.float32(s["mint"]), np.float32(s["maxt"])
            assert mint >= 0  # rays from outside the scene: no inside segments
            d = np.float32(u * (np.float64(maxt - mint) + 1e-3) + np.float64(mint))
            t = d if t is None else np.float32(np.float32(0.5) * np.float32(t + d))
        if t is None:
            assert nb.hit_kind(hits[i:i + 1])[0] == 0
            continue
        if mode == oracle.FP64:
            assert hits["t"][i] == t, (i, hits["t"][i], t)
        else:
            assert abs(hits["t"][i] - t) <= 1e-6 * t, (i, hits["t"][i], t)
        folded += len(segs) > 1
    assert folded > 20
    # equal normals stay the same normal
    hit = nb.hit_kind(hits) == 1
    v = sig(LOGITS[2:5].astype(np.float64)) * 2 - 1
    n = v / np.linalg.norm(v)
    assert np.allclose(hits["normal_or_uv"][hit], n, atol=1e-6)
    # without interpolation: the minimum over segments
    plain = constant_network(nb, oracle, nb.network_config(), LOGITS, scene["bmin"], scene["bmax"])
    ph = plain.neural_query(scene["nbvh"], scene["rays"], mode=mode)
    assert np.all(ph["t"][hit] <= hits["t"][hit]) and np.any(ph["t"][hit] < hits["t"][hit])
