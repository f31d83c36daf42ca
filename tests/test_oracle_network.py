"""Pins of the oracle's encoding, hash grid and MLP.

tiny-cuda-nn is absent from the reference tree, so its exact grid convention is restated from
its public GridEncoding (scale = b^l * N_min - 1, resolution ceil(scale) + 1, pos = scale*x + 0.5)
and checked here against independent numpy restatements and mathematical properties:
textbook trilinear interpolation of a dense level, the coherent prime hash, FP64 GEMMs.
"""
import numpy as np
import pytest

from conftest import camera_rays, random_rays

P1, P2 = np.uint64(2654435761), np.uint64(805459861)


def small_cfg(nb, L=2, F=2, log2T=12, base=4, scale=2.0, dist=1):
    return nb.network_config(n_levels=L, n_features_per_level=F, log2_hashmap_size=log2T, base_resolution=base,
                             per_level_scale=scale, segment_distribution=dist)


def make_net(nb, orc, cfg, seed=3, table_range=1.0):
    tables, mlp = nb.network_init_params(cfg, seed, table_range)
    return orc.Net(cfg, tables, mlp, (-1, -1, -1), (1, 1, 1)), tables, mlp


def test_level_table_matches_host_library(nb, orc):
    cfg = nb.network_config()
    net, tables, _ = make_net(nb, orc, cfg, table_range=1e-4)
    lv = net.levels()
    assert lv["n_params"] == nb.network_param_counts(cfg)[0] == len(tables)
    # Benchmark grid: 16 levels, base 16, scale 1.5, 2^19 entries: levels 0..3 dense.
    assert list(lv["dense"]) == [1] * 4 + [0] * 12
    assert list(lv["res"][:4]) == [16, 24, 36, 54]
    assert list(lv["size"][:4]) == [4096, 13824, 46656, 157464] and all(lv["size"][4:] == 2 ** 19)
    assert lv["n_params"] == 2 * (4096 + 13824 + 46656 + 157464 + 12 * 2 ** 19)


def numpy_grid_point(x, lv, tables, F, l):
    """Independent restatement: trilinear interpolation over the 8 corners."""
    scale, res, size, off, dense = (lv[k][l] for k in ("scale", "res", "size", "offset", "dense"))
    pos = np.float64(np.float32(scale)) * x + 0.5
    g = np.floor(pos)
    f = pos - g
    g = g.astype(np.int64)
    acc = np.zeros(F)
    for c in range(8):
        bits = [(c >> d) & 1 for d in range(3)]
        w = np.prod([f[d] if bits[d] else 1 - f[d] for d in range(3)])
        gc = [np.uint64((g[d] + bits[d]) & 0xFFFFFFFF) for d in range(3)]
        if dense:
            idx = int((gc[0] + gc[1] * np.uint64(res) + gc[2] * np.uint64(res) ** 2) & np.uint64(0xFFFFFFFF)) % size
        else:
            h = (gc[0] ^ ((gc[1] * P1) & np.uint64(0xFFFFFFFF)) ^ ((gc[2] * P2) & np.uint64(0xFFFFFFFF)))
            idx = int(h) % size
        acc += w * tables[(off + idx) * F:(off + idx) * F + F].astype(np.float64)
    return acc


def test_dense_level_is_textbook_trilinear(nb, orc):
    # One dense level (res^3 <= T): a linear field stored at the vertices is reproduced exactly
    # by trilinear interpolation (in FP64), which pins cell/corner/weight conventions.
    cfg = small_cfg(nb, L=1, F=2, log2T=14, base=9)
    net, tables, _ = make_net(nb, orc, cfg)
    lv = net.levels()
    assert lv["dense"][0] == 1 and lv["res"][0] == 9
    res, scale = 9, float(lv["scale"][0])
    grid = np.stack(np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij"), -1)
    # vertex (i, j, k) -> index i + res*j + res^2*k; value = linear function of the vertex position
    coeff = np.array([0.3, -0.2, 0.7])
    vals = (grid.reshape(-1, 3) - 0.5) / scale @ coeff + 0.1
    t = np.zeros_like(tables)
    idx = (grid[..., 0] + res * grid[..., 1] + res * res * grid[..., 2]).reshape(-1)
    t[idx * 2] = vals
    t[idx * 2 + 1] = 2 * vals
    net2 = orc.Net(cfg, t, net.mlp, (-1, -1, -1), (1, 1, 1))
    # pos = scale*x + 0.5; cells whose +1 corner stays inside the res^3 vertices (tcnn lets the
    # last half cell reach coordinate res, which aliases into the next row).
    x = np.random.default_rng(0).uniform(0, (res - 1.5) / scale, (200, 3))
    feats = net2.grid(x, mode=orc.FP64)
    expect = x @ coeff + 0.1
    assert np.allclose(feats[:, 0], expect, atol=1e-6)
    assert np.allclose(feats[:, 1], 2 * expect, atol=1e-6)


def test_grid_matches_independent_numpy(nb, orc):
    cfg = small_cfg(nb, L=6, F=2, log2T=10, base=5, scale=1.7)
    net, tables, _ = make_net(nb, orc, cfg)
    lv = net.levels()
    assert lv["dense"][0] == 1 and lv["dense"][-1] == 0
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 1, (100, 3)), rng.uniform(-0.05, 1.05, (30, 3))])
    feats = net.grid(x, mode=orc.FP64)
    for i in range(len(x)):
        for l in range(6):
            np.testing.assert_allclose(feats[i, 2 * l:2 * l + 2], numpy_grid_point(x[i], lv, tables, 2, l),
                                       rtol=1e-12, atol=1e-12)


def test_grid_fp32_close_to_fp64(nb, orc):
    cfg = nb.network_config(log2_hashmap_size=14)
    net, _, _ = make_net(nb, orc, cfg, table_range=1e-4)
    x = np.random.default_rng(6).uniform(0, 1, (500, 3)).astype(np.float32)
    f32 = net.grid(x, mode=orc.FP32).astype(np.float64)
    f64 = net.grid(x.astype(np.float64), mode=orc.FP64)
    # FP32 pos = scale*x + 0.5 carries an absolute error ~ scale * 2^-24, so the corner weights of
    # level l are off by up to ~scale_l * 2^-23; features (8 corners of |v| <= 1e-4) by at most
    # 8 * 1e-4 * scale_l * 2^-23 (rarely more: a rounding that crosses a cell boundary).
    scales = np.repeat(net.levels()["scale"].astype(np.float64), 2)
    bound = 8 * 1e-4 * scales * 2.0 ** -23 + 1e-12
    within = np.abs(f32 - f64) <= bound[None, :]
    assert within.mean() > 0.999
    assert within[:, :8].all()  # coarse levels


def numpy_mlp(mlp, x, K0, H=64, NH=4, O=8):
    off = 0
    a = x.astype(np.float64)
    n_in = K0
    for layer in range(NH + 1):
        n_out = O if layer == NH else H
        W = mlp[off:off + n_out * n_in].astype(np.float64).reshape(n_out, n_in)
        off += n_out * n_in
        a = a @ W.T
        if layer < NH:
            a = np.maximum(a, 0)
        n_in = H
    return a


def test_mlp_matches_numpy_fp64(nb, orc):
    cfg = nb.network_config(log2_hashmap_size=14)
    net, tables, mlp = make_net(nb, orc, cfg, table_range=0.5)
    rng = np.random.default_rng(8)
    inputs = rng.uniform(0.02, 0.98, (64, 9)).astype(np.float32)
    outs, logits = net.network(inputs, mode=orc.FP64)
    feats = np.concatenate([net.grid(inputs[:, 3 * k:3 * k + 3].astype(np.float64), mode=orc.FP64)
                            for k in range(3)], axis=1)
    z = numpy_mlp(mlp, feats, 96)
    np.testing.assert_allclose(logits, z, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(outs, 1 / (1 + np.exp(-z)), rtol=1e-12)
    # FP32 mode agrees with FP64 up to the FP32 grid-position error of the fine levels
    # (weights off by ~scale * 2^-23, see test_grid_fp32_close_to_fp64).
    o32, z32 = net.network(inputs, mode=orc.FP32)
    np.testing.assert_allclose(z32, z, rtol=0, atol=2e-3 * np.abs(z).max())


def test_mlp_plants_detected(nb, orc):
    # A transposed weight matrix or a dropped ReLU changes the logits: the comparison has teeth.
    cfg = nb.network_config(log2_hashmap_size=14)
    net, tables, mlp = make_net(nb, orc, cfg, table_range=0.5)
    inputs = np.random.default_rng(9).uniform(0.1, 0.9, (16, 9)).astype(np.float32)
    _, z = net.network(inputs, mode=orc.FP64)
    feats = np.concatenate([net.grid(inputs[:, 3 * k:3 * k + 3].astype(np.float64), mode=orc.FP64)
                            for k in range(3)], axis=1)
    bad = mlp.copy()
    W2 = bad[96 * 64:96 * 64 + 4096].reshape(64, 64)
    bad[96 * 64:96 * 64 + 4096] = W2.T.reshape(-1)
    assert not np.allclose(numpy_mlp(bad, feats, 96), z, rtol=1e-6)


def test_encode_closed_form(nb, orc):
    cfg = nb.network_config(log2_hashmap_size=14)
    tables, mlp = nb.network_init_params(cfg, 1)
    bmin, bmax = np.float32([-1.2, -0.9, -1.1]), np.float32([1.1, 1.3, 0.8])
    net = orc.Net(cfg, tables, mlp, bmin, bmax)
    rays = random_rays(50, 12)
    mint = np.random.default_rng(1).uniform(-0.5, 2, 50).astype(np.float32)
    maxt = (np.abs(mint) + np.random.default_rng(2).uniform(0.1, 1, 50)).astype(np.float32)
    enc = net.encode(rays, mint, maxt, mode=orc.FP64)
    dil = np.float64(np.float32(np.float32(1 + np.float32(1e-4)) + np.float32(1e-5)))
    c = (bmax.astype(np.float64) + bmin) / 2
    ext = (bmax.astype(np.float64) - bmin) * dil
    for i in range(50):
        a, b = (maxt[i], mint[i]) if mint[i] < 0 else (mint[i], maxt[i])  # negative mint: swapped
        po = rays["o"][i].astype(np.float64) + np.float64(a) * rays["d"][i]
        pi = rays["o"][i].astype(np.float64) + np.float64(b) * rays["d"][i]
        eo, ei = (po - c) / ext + 0.5, (pi - c) / ext + 0.5
        for k, t in enumerate((0.0, 0.5, 1.0)):  # EquidistantExterior, 3 points
            np.testing.assert_allclose(enc[i, 3 * k:3 * k + 3], (1 - t) * eo + t * ei, rtol=1e-6, atol=1e-6)
    # EquidistantInterior: t_k = (k + 0.5) / 3
    cfg2 = nb.network_config(log2_hashmap_size=14, segment_distribution=0)
    net2 = orc.Net(cfg2, tables, mlp, bmin, bmax)
    e2 = net2.encode(rays, mint, maxt, mode=orc.FP64)
    np.testing.assert_allclose(e2[:, 3:6], 0.5 * (enc[:, 0:3] + enc[:, 6:9]), rtol=1e-6, atol=1e-7)


def test_neural_query_fp32_close_to_fp64(nb, orc, small_scene):
    v, t, nodes, prims, nbn = small_scene
    cfg = nb.network_config(log2_hashmap_size=14)
    tables, mlp = nb.network_init_params(cfg, 5)
    net = orc.Net(cfg, tables, mlp, nodes[0]["aabb_min"], nodes[0]["aabb_max"])
    _, rays = camera_rays(nb, orc, 64, 64)
    rays = np.concatenate([rays, random_rays(1000, 3, inside_frac=0.3)])
    h32 = net.neural_query(nbn, rays, mode=orc.FP32)
    h64, log = net.neural_query(nbn, rays, mode=orc.FP64, log_segments=True)
    k32, k64 = h32["info"] & 0xFF, h64["info"] & 0xFF
    assert (k64 == 1).mean() > 0.3
    assert (k32 != k64).mean() < 0.005
    both = (k32 == 1) & (k64 == 1) & (h32["node_id"] == h64["node_id"])
    np.testing.assert_allclose(h32["t"][both], h64["t"][both], rtol=1e-4, atol=1e-5)
    # Segment log is consistent with the hit records.
    assert len(log) == int((h64["info"] >> 16).sum())
    # Inside-node rays exist in the mix (negative mint path).
    assert (log["mint"] < 0).sum() > 0
