"""Pins of the oracle's geometry: slab test, node dilation, N-BVH traversal (leaf sets against a
brute-force cut of the tree), the shipped chess .nbvh, and the triangle BVH against brute force."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, camera_rays, random_rays

CHESS = os.path.join(GOLDEN, "chess_nodes_32940_hash_18_lvl_8.nbvh")
FLT_MAX = np.float32(3.4028234663852886e38)


def np_slab(bmin, bmax, o, d):
    """Majercik slab test in float32 numpy with glm::min/max NaN semantics."""
    with np.errstate(all="ignore"):
        t0 = (np.float32(bmin) - o) / d
        t1 = (np.float32(bmax) - o) / d
    gmin = lambda x, y: np.where(y < x, y, x)
    gmax = lambda x, y: np.where(x < y, y, x)
    s, b = gmin(t0, t1), gmax(t0, t1)
    tn = gmax(gmax(s[..., 0], s[..., 1]), s[..., 2])
    tf = gmin(gmin(b[..., 0], b[..., 1]), b[..., 2])
    return tn, tf


def test_slab_matches_numpy_including_axis_aligned(orc):
    rng = np.random.default_rng(1)
    rays = random_rays(300, 2)
    # Axis-aligned directions (zero components: +-inf and 0/0 = NaN slabs) and origins on planes.
    rays["d"][:60] = np.float32([[1, 0, 0], [0, -1, 0], [0, 0, 1], [1, 1, 0], [0, 1, -1], [-1, 0, 0]] * 10)
    rays["o"][30:60, 0] = np.float32(-0.5)
    for i in range(len(rays)):
        bmin = rng.uniform(-1, 0, 3).astype(np.float32)
        bmax = (bmin + rng.uniform(0.1, 1.5, 3)).astype(np.float32)
        if i < 60:
            bmin[0] = np.float32(-0.5)
        hit, tn, tf = orc.slab(bmin, bmax, rays[i])
        etn, etf = np_slab(bmin, bmax, rays["o"][i], rays["d"][i])
        assert np.array_equal(np.float32(tn), etn, equal_nan=True)
        assert np.array_equal(np.float32(tf), etf, equal_nan=True)
        assert hit == bool(etn <= etf)


def test_slab_interval_brute_force(orc):
    # The returned [tn, tf] contains exactly the sampled ray parameters whose points lie in the box.
    rng = np.random.default_rng(3)
    for _ in range(200):
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ray = np.zeros(1, orc.RAY_DTYPE)[0]
        ray["o"], ray["d"] = o, d
        bmin = rng.uniform(-1, 0, 3).astype(np.float32)
        bmax = (bmin + rng.uniform(0.2, 1.2, 3)).astype(np.float32)
        hit, tn, tf = orc.slab(bmin, bmax, ray)
        ts = np.linspace(-8, 8, 4001)
        pts = o[None, :].astype(np.float64) + ts[:, None] * d[None, :].astype(np.float64)
        inside = np.all((pts >= bmin - 1e-6) & (pts <= bmax + 1e-6), axis=1)
        strict = np.all((pts > bmin + 1e-4) & (pts < bmax - 1e-4), axis=1)
        if strict.any():
            assert hit and tn <= ts[strict].min() + 1e-5 and tf >= ts[strict].max() - 1e-5
        if hit and tf - tn > 0.01:
            mid = 0.5 * (tn + tf)
            assert inside[np.argmin(np.abs(ts - mid))]
        if not inside.any():
            assert (not hit) or (tf - tn) < 1e-3


def test_dilation_closed_form(orc, nb):
    node = np.zeros(1, nb.NODE_DTYPE)[0]
    node["aabb_min"] = np.float32([-1.0, 0.25, 3.0])
    node["aabb_max"] = np.float32([1.0, 0.5, 7.0])
    a, b = orc.dilate(node, 1e-5)
    ext = node["aabb_max"] - node["aabb_min"]
    c = node["aabb_min"] + ext * np.float32(0.5)
    e2 = ext * np.float32(1.0 + np.float32(1e-5))
    assert np.array_equal(a, c - e2 * np.float32(0.5)) and np.array_equal(b, c + e2 * np.float32(0.5))
    # Dilation grows the box symmetrically by ext * 0.5e-5 (up to rounding).
    assert np.allclose(node["aabb_min"] - a, ext * 0.5e-5, rtol=1e-2, atol=1e-7)


def brute_leaf_set(orc, nodes, ray, lod):
    """Leaves of the cut at `lod` whose dilated box passes the inference acceptance rule; the cut
    is found by a plain recursive walk over the tree (no stack discipline, no culling)."""
    out = set()
    stack = [0]
    while stack:
        i = stack.pop()
        nd = nodes[i]
        pc, left = int(nd["prim_count"]), int(nd["left_or_prim_offset"])
        leaf = (0 < pc <= 2 + lod) or left == 0xFFFFFFFF
        a, b = orc.dilate(nd, 1e-5)
        hit, tn, tf = orc.slab(a, b, ray)
        if not leaf:
            if hit:
                stack += [left, left + 1]
            continue
        if not hit:
            continue
        ok = tn < FLT_MAX and ((tf > np.float32(1e-3)) if tn < 0 else (tn > np.float32(1e-4)))
        if ok:
            out.add(i)
    return out


def check_leaf_sets(orc, nodes, rays, lod):
    n_nonempty = 0
    for r in rays:
        seq = orc.leaf_sequence(nodes, r, lod=lod)
        assert seq is not None, "stack overflow"
        ids = list(seq[0])
        assert len(ids) == len(set(ids)), "a leaf visited twice"
        assert set(ids) == brute_leaf_set(orc, nodes, r, lod)
        n_nonempty += len(ids) > 0
        # Segments are the dilated slab interval of the visited leaf.
        for node, mn, mx in zip(*seq):
            a, b = orc.dilate(nodes[node], 1e-5)
            _, tn, tf = orc.slab(a, b, r)
            assert mn == tn and mx == tf
    return n_nonempty


def test_leaf_sets_bench_cut(nb, orc, small_scene):
    v, t, nodes, prims, nbn = small_scene
    rays = random_rays(150, 5, inside_frac=0.3)
    assert check_leaf_sets(orc, nbn, rays, 0) > 60


def test_chess_fixture_format(nb):
    data = open(CHESS, "rb").read()
    assert len(data) == 1009984
    nodes = nb.nbvh_parse(data)
    assert len(nodes) == 25249
    pc = nodes["prim_count"]
    hist = {int(k): int(c) for k, c in zip(*np.unique(pc, return_counts=True))}
    assert hist == {0: 12624, 1: 106, 8: 12519}
    leaves = nodes["left_or_prim_offset"] == 0xFFFFFFFF
    assert leaves.sum() == 12625
    assert np.allclose(nodes[0]["aabb_min"], [-0.2457, 0.0149, -0.2491], atol=1e-4)
    assert np.allclose(nodes[0]["aabb_max"], [0.2469, 0.169, 0.2481], atol=1e-4)
    # Serialisation round trip is byte-exact.
    assert nb.nbvh_serialize(nodes) == data


def test_leaf_sets_chess_fixture(nb, orc):
    nodes = nb.nbvh_parse(open(CHESS, "rb").read())
    lo, hi = nodes[0]["aabb_min"], nodes[0]["aabb_max"]
    rng = np.random.default_rng(11)
    rays = np.zeros(60, orc.RAY_DTYPE)
    o = rng.uniform(lo - 0.3, hi + 0.3, (60, 3))
    tgt = rng.uniform(lo, hi, (60, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"], rays["d"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, FLT_MAX
    for lod in (0, 6):
        assert check_leaf_sets(orc, nodes, rays, lod) > 20


def test_triangle_bvh_matches_brute_force(nb, orc, small_scene):
    v, t, nodes, prims, nbn = small_scene
    _, rays = camera_rays(nb, orc, 48, 48)
    rays = np.concatenate([rays, random_rays(400, 9, inside_frac=0.5)])
    h = orc.tri_query(v, t, nodes, prims, rays)
    b = orc.tri_brute(v, t, rays)
    assert np.array_equal(h["info"] & 0xFF, b["info"] & 0xFF)
    assert np.array_equal(h["t"], b["t"])
    assert np.array_equal(h["prim_id"], b["prim_id"])
    assert (h["info"] & 0xFF == 2).mean() > 0.3
    # Any-hit finds a hit exactly when closest-hit does.
    a = orc.tri_query(v, t, nodes, prims, rays, any_hit=True)
    assert np.array_equal(a["info"] & 0xFF, h["info"] & 0xFF)


def test_watertight_shared_edges(nb, orc):
    # Rays through shared vertices/edges of a closed mesh never slip through (watertightness):
    # every ray from outside aimed at the centre of a closed sphere hits.
    v, t = nb.icosphere(2, seed=0, amplitude=0.0)
    nodes, prims = nb.bvh_build(v, t)
    rays = np.zeros(len(v), orc.RAY_DTYPE)
    rays["o"] = (v * 3.0).astype(np.float32)
    d = -v / np.linalg.norm(v, axis=1, keepdims=True)
    rays["d"] = d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, FLT_MAX
    h = orc.tri_query(v, t, nodes, prims, rays)
    assert np.all(h["info"] & 0xFF == 2)


def test_scatter_diffuse_is_cosine_distributed(orc):
    """The bounce restatement: unit directions in the hemisphere of the normal turned toward the
    incoming ray, cosine-weighted (E[cos] = 2/3, E[cos^2] = 1/2 for p(w) = cos/pi; a uniform
    hemisphere would give 1/2 and 1/3), origin = hit point + 1e-4 * normal, hit-less rays inactive."""
    n = 200000
    rng = np.random.default_rng(4)
    rays = np.zeros(n, orc.RAY_DTYPE)
    rays["o"] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    hits = np.zeros(n, orc.HIT_DTYPE)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    hits["normal_or_uv"] = nrm.astype(np.float32)
    hits["t"] = rng.uniform(0.1, 2.0, n).astype(np.float32)
    hits["info"] = 1
    hits["info"][:1000] = 0  # misses
    out = orc.scatter_diffuse(rays, hits, sample=3, bounce=1)
    assert np.all(out["tmax"][:1000] == -1) and np.all(out["tmax"][1000:] > 1e38)
    live = slice(1000, None)
    wi = out["d"][live].astype(np.float64)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-5)
    nf = np.asarray(hits["normal_or_uv"][live], np.float64)
    facing = np.einsum("ij,ij->i", nf, rays["d"][live].astype(np.float64)) > 0
    nf[facing] *= -1  # turned toward the incoming ray
    cos = np.einsum("ij,ij->i", wi, nf)
    assert cos.min() > -1e-5
    assert abs(cos.mean() - 2 / 3) < 0.004 and abs((cos ** 2).mean() - 0.5) < 0.004
    p = rays["o"][live].astype(np.float64) + hits["t"][live, None] * rays["d"][live].astype(np.float64)
    assert np.allclose(out["o"][live], p + 1e-4 * nf, atol=2e-6)
    # Deterministic in (sample, bounce, ray); a different bounce draws different directions.
    again = orc.scatter_diffuse(rays, hits, sample=3, bounce=1)
    other = orc.scatter_diffuse(rays, hits, sample=3, bounce=2)
    assert np.array_equal(again.view(np.uint32), out.view(np.uint32))
    assert np.mean(np.all(other["d"][live] == out["d"][live], axis=1)) < 1e-3


def test_scatter_diffuse_triangle_normal(nb, orc):
    v, t = nb.icosphere(2, amplitude=0.0)
    rays = np.zeros(len(t), orc.RAY_DTYPE)
    cen = v[t].mean(axis=1)
    rays["o"] = (cen * 3).astype(np.float32)
    rays["d"] = (-cen / np.linalg.norm(cen, axis=1, keepdims=True)).astype(np.float32)
    hits = np.zeros(len(t), orc.HIT_DTYPE)
    hits["info"] = 2
    hits["prim_id"] = np.arange(len(t))
    hits["t"] = np.linalg.norm(cen * 2, axis=1).astype(np.float32)
    out = orc.scatter_diffuse(rays, hits, verts=v, tris=t)
    # Outward geometric normal (counter-clockwise faces), facing the incoming ray from outside.
    g = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    assert np.all(np.einsum("ij,ij->i", out["d"], g) > -1e-5)
    assert np.allclose(out["o"], cen + 1e-4 * g, atol=1e-5)


def test_segment_labels_restatement(nb, orc, small_scene):
    """Labels of real segments (the oracle's own segment log): visibility agrees with brute-force
    triangle intersection inside [mint, maxt), depth in [0, 1) and equal to the closed form, normal
    (n+1)/2 of the hit triangle's face normal, albedo the default 0.9, zeros without a hit."""
    v, t, nodes, prims, nbn = small_scene
    cfg = nb.network_config(n_levels=2, log2_hashmap_size=10)
    tables, mlp = nb.network_init_params(cfg, 1, 1e-4)
    lo, hi = nodes[0]["aabb_min"], nodes[0]["aabb_max"]
    net = orc.Net(cfg, tables, mlp, lo, hi)
    rays = random_rays(400, 17, inside_frac=0.3)
    _, log = net.neural_query(nbn, rays, threshold=0.999999, mode=orc.FP32, log_segments=True)  # every leaf
    r = rays[log["ray"]]
    mint = np.maximum(log["mint"], 0).astype(np.float32)
    maxt = log["maxt"]
    lab = orc.segment_labels(v, t, nodes, prims, r, mint, maxt)
    rr = np.array(r, copy=True)
    rr["tmin"], rr["tmax"] = mint, maxt
    brute = orc.tri_brute(v, t, rr)  # closest hit at t >= tmin, any distance
    hit = ((brute["info"] & 0xFF) == 2) & (brute["t"] < maxt)
    assert hit.mean() > 0.1 and (~hit).mean() > 0.02
    assert np.array_equal(lab[:, 0] == 1, hit)
    assert np.all(lab[~hit] == 0)
    d = lab[hit, 1]
    assert np.all((d >= 0) & (d < 1))
    exp_d = (brute["t"][hit] - mint[hit]) / ((maxt[hit] - mint[hit]).astype(np.float64) + 1e-3)
    assert np.allclose(d, exp_d, rtol=1e-6, atol=1e-7)
    tri = t[brute["prim_id"][hit]].astype(np.int64)
    n = np.cross(v[tri[:, 1]] - v[tri[:, 0]], v[tri[:, 2]] - v[tri[:, 0]]).astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.allclose(lab[hit, 2:5], (n + 1) / 2, atol=1e-6)
    assert np.all(lab[hit, 5:8] == np.float32(0.9))
