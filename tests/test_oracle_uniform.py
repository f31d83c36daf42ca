"""The uniform segment distribution of the oracle (neural_bvh_module_kernels.cu:53-67): for segment
i the n_points parameters t_k are successive lcg_randomf draws of the stream
get_lcg_rng(segment_seed, 0, i), and the points are (1 - t) eo + t ei in FP32 between the
segment's encoded end points. The stream is restated here from the LCG's closed form (its seeding
is pinned by tests/test_oracle_rng.py); the end points come from the exterior distribution, whose
first and last points are exactly eo and ei."""
import numpy as np

import oracle


def segments(nb, n=300, seed=4):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, nb.RAY_DTYPE)
    o = rng.uniform(-1.5, 1.5, (n, 3))
    d = rng.normal(size=(n, 3))
    rays["o"] = o.astype(np.float32)
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(3.4e38)
    mint = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    maxt = (mint + rng.uniform(0.05, 1.0, n)).astype(np.float32)
    return rays, mint, maxt


def net(nb, dist, n_points=3):
    cfg = nb.network_config(segment_distribution=dist, n_points=n_points)
    probe = oracle.Net(cfg, np.zeros(1, np.float32), np.zeros(1, np.float32), (-1.6,) * 3, (1.6,) * 3)
    n_params = int(probe.levels()["n_params"])
    tables = np.zeros(n_params * cfg.grid.n_features_per_level, np.float32)
    mlp = np.zeros(64 * 96 + 3 * 64 * 64 + 8 * 64, np.float32)
    return oracle.Net(cfg, tables, mlp, (-1.6, -1.5, -1.4), (1.6, 1.5, 1.4))


def draws(orc, seed, i, k):
    """k successive lcg_randomf draws of get_lcg_rng(seed, 0, i): state = 1664525 state + 1013904223
    (mod 2^32), value = float32(state) * 2^-32."""
    st = int(orc.oracle_lcg_seed(seed, 0, i))
    out = []
    for _ in range(k):
        st = (1664525 * st + 1013904223) & 0xFFFFFFFF
        out.append(np.float32(np.float32(st) * np.float32(2.0 ** -32)))
    return out


def test_uniform_points_from_the_ray_stream(nb):
    orc = oracle.lib()
    rays, mint, maxt = segments(nb)
    ends = net(nb, nb.SEGMENT_EQUIDISTANT_EXTERIOR).encode(rays, mint, maxt).reshape(-1, 3, 3)
    eo, ei = ends[:, 0], ends[:, 2]
    uni = net(nb, nb.SEGMENT_UNIFORM)
    for seed in (0, 12345):
        pts = uni.encode(rays, mint, maxt, segment_seed=seed).reshape(-1, 3, 3)
        for i in range(len(rays)):
            for k, t in enumerate(draws(orc, seed, i, 3)):
                w = np.float32(np.float32(1.0) - t)
                exp = (w * eo[i]).astype(np.float32) + (t * ei[i]).astype(np.float32)
                assert np.array_equal(pts[i, k], exp.astype(np.float32)), (seed, i, k)
    # other seeds, other points; the FP64 mode draws the same parameters
    a = uni.encode(rays, mint, maxt, segment_seed=1)
    b = uni.encode(rays, mint, maxt, segment_seed=2)
    assert not np.array_equal(a, b)
    c = uni.encode(rays, mint, maxt, mode=oracle.FP64, segment_seed=1)
    assert np.allclose(a, c, atol=1e-6)
