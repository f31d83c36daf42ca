"""GPU parity: the HIP kernels, called through the C ABI, against the CPU oracle.

Bars: bit-exact for rays, triangle ids/t, N-BVH leaf sequences and the whole FP32 neural query
(the FP32 path evaluates the oracle's FP32 operation sequence); FP16/BF16 network outputs within
stated tolerances of the FP64 oracle, visibility flips <= 0.5 %.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, bench_scene, camera_rays, random_rays

pytestmark = pytest.mark.gpu

CHESS = os.path.join(GOLDEN, "chess_nodes_32940_hash_18_lvl_8.nbvh")


def hits_equal(a, b):
    """Bitwise equality of hit records (NaN-safe)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    w = a.dtype.itemsize // 4
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint32).reshape(len(a), w),
                                                 b.view(np.uint32).reshape(len(b), w))


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def bench(nb, orc, ctx):
    v, t, nodes, prims, nbn = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    cfg = nb.network_config()  # 16 levels x 2 features, 2^19, base 16, scale 1.5
    scene.init_network(cfg, seed=1, table_range=1e-4)
    cfg2, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    net = orc.Net(cfg2, tables, mlp, bmin, bmax)
    yield dict(v=v, t=t, nodes=nodes, prims=prims, nbvh=scene.nbvh(), scene=scene, net=net,
               module=nb.NeuralBVHModule(ctx, scene))
    scene.close()


def test_scene_matches_host_builders(nb, bench):
    nodes, prims = bench["scene"].bvh()
    assert np.array_equal(nodes.view(np.uint32), bench["nodes"].view(np.uint32))
    assert np.array_equal(prims, bench["prims"])
    assert np.array_equal(bench["nbvh"].view(np.uint32), nb.nbvh_cut(nodes, 256).view(np.uint32))


def test_raygen_bit_exact(nb, orc, ctx):
    from nbvh_amd import hip
    w, h = 512, 384
    cam, expect = camera_rays(nb, orc, w, h, sample=3, frame=2)
    d = hip.Buffer(w * h * 32)
    ctx.raygen(cam, w, h, d.ptr, sample=3, frame=2)
    hip.synchronize()
    got = d.download(orc.RAY_DTYPE)
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))


def test_device_query_on_stream_matches_host_query(nb, orc, bench):
    from nbvh_amd import hip
    _, rays = camera_rays(nb, orc, 64, 64)
    exp = bench["module"].neural_bvh_traversal(rays, precision=nb.PREC_FP16)
    s = hip.Stream()
    dr, dh = hip.Buffer.from_array(rays), hip.Buffer(len(rays) * 48)
    p = bench["module"].params(precision=nb.PREC_FP16)
    bench["module"].query_device(p, dr.ptr, len(rays), dh.ptr, stream=s.handle)
    s.synchronize()
    got = dh.download(nb.HIT_DTYPE)
    assert hits_equal(got, exp)
    s.destroy()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_tiled_claim_order_same_results(nb, orc, bench, precision):
    # The image-width hint only changes which wave traces which ray.
    prec = {"fp32": nb.PREC_FP32, "fp16": nb.PREC_FP16}[precision]
    _, rays = camera_rays(nb, orc, 96, 64)
    a = bench["module"].neural_bvh_traversal(rays, precision=prec)
    b = bench["module"].neural_bvh_traversal(rays, precision=prec, image_width=96)
    assert hits_equal(a, b)
    c = bench["module"].neural_bvh_traversal(rays[:96 * 60], precision=prec, image_width=96)  # 60 % 8: ignored
    assert hits_equal(a[:96 * 60], c)


@pytest.mark.parametrize("any_hit", [False, True])
def test_triangle_query_bit_exact(nb, orc, bench, any_hit):
    _, rays = camera_rays(nb, orc, 256, 256)
    rays = np.concatenate([rays, random_rays(20000, 4, inside_frac=0.3)])
    kind = nb.QUERY_TRIANGLE_ANY if any_hit else nb.QUERY_TRIANGLE_CLOSEST
    got = bench["module"].neural_bvh_traversal(rays, kind=kind)
    exp = orc.tri_query(bench["v"], bench["t"], bench["nodes"], bench["prims"], rays, any_hit=any_hit)
    if not any_hit:
        assert hits_equal(got, exp)
    else:
        assert np.array_equal(got["info"], exp["info"])
    assert (nb.hit_kind(got) == nb.HIT_TRIANGLE).mean() > 0.3


def test_network_fp32_bit_exact(nb, orc, bench):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(0, 1, (3000, 9)), rng.uniform(-0.2, 1.2, (500, 9))]).astype(np.float32)
    out, logits = bench["module"].network_inference(x, precision=nb.PREC_FP32)
    o32, z32 = bench["net"].network(x, mode=orc.FP32)
    assert np.array_equal(logits, z32.astype(np.float32))
    assert np.array_equal(out, o32.astype(np.float32))


@pytest.mark.parametrize("any_hit", [False, True])
def test_neural_query_fp32_bit_exact(nb, orc, bench, any_hit):
    _, rays = camera_rays(nb, orc, 128, 128)
    rays = np.concatenate([rays, random_rays(6000, 7, inside_frac=0.4)])
    kind = nb.QUERY_NEURAL_ANY if any_hit else nb.QUERY_NEURAL_CLOSEST
    got = bench["module"].neural_bvh_traversal(rays, kind=kind, precision=nb.PREC_FP32)
    exp = bench["net"].neural_query(bench["nbvh"], rays, any_hit=any_hit, mode=orc.FP32)
    assert (nb.hit_kind(exp) == nb.HIT_NEURAL).mean() > 0.3
    assert np.array_equal(got["info"], exp["info"])
    assert hits_equal(got, exp)


@pytest.mark.parametrize("world_t,linear,interp", [(True, False, False), (False, True, True), (True, True, True)])
def test_decode_options_bit_exact(nb, orc, ctx, bench, world_t, linear, interp):
    """The network's decode options (WorldT depth, linear depth activation, depth interpolation
    within 2e-2) on the benchmark scene: FP32 closest and any hit bit-exact with the oracle,
    the wavefront-round schedule identical to the persistent kernel, FP16 hit/miss within 0.5 %
    of the FP64 oracle."""
    cfg = nb.network_config(depth_mode=nb.DEPTH_WORLD_T if world_t else nb.DEPTH_LOCAL_T,
                            depth_linear_activation=linear, interpolate_depth=interp, depth_interpolation_eps=2e-2)
    scene = nb.Scene(ctx, bench["v"], bench["t"])
    scene.build_nbvh(256)
    scene.init_network(cfg, seed=3, table_range=1e-1)
    cfg2, tables, mlp = scene.network()
    assert (cfg2.depth_mode, cfg2.depth_linear_activation, cfg2.interpolate_depth) == (int(world_t), int(linear),
                                                                                       int(interp))
    bmin, bmax = scene.bbox()
    net = orc.Net(cfg2, tables, mlp, bmin, bmax)
    mod = nb.NeuralBVHModule(ctx, scene)
    _, rays = camera_rays(nb, orc, 96, 96)
    rays = np.concatenate([rays, random_rays(5000, 17, inside_frac=0.4)])
    for any_hit in (False, True):
        kind = nb.QUERY_NEURAL_ANY if any_hit else nb.QUERY_NEURAL_CLOSEST
        got = mod.neural_bvh_traversal(rays, kind=kind, precision=nb.PREC_FP32)
        exp = net.neural_query(scene.nbvh(), rays, any_hit=any_hit, mode=orc.FP32)
        assert 0.1 < (nb.hit_kind(exp) == nb.HIT_NEURAL).mean() < 0.95
        assert hits_equal(got, exp), any_hit
        wf = mod.neural_bvh_traversal(rays, kind=kind, precision=nb.PREC_FP32, wavefront=True)
        assert hits_equal(wf, got), any_hit
    g16 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16)
    e64 = net.neural_query(scene.nbvh(), rays, mode=orc.FP64)
    assert (nb.hit_kind(g16) != nb.hit_kind(e64)).mean() <= 0.005
    if world_t or linear:  # the trainer's loss is for the default decoding only
        with pytest.raises(nb.NbvhError):
            nb.Trainer(ctx, scene, max_batch=64)
    scene.close()


def test_uniform_segment_distribution(nb, orc, ctx, bench):
    """The uniform segment distribution (random segment points at inference,
    neural_bvh_module_kernels.cu:53-67): FP32 closest and any hit bit-exact with the oracle for two
    seeds (each ray's stream seeded by its index, n_points draws per evaluated segment), other
    seeds other hits, FP16 hit/miss within 0.5 % of the FP64 oracle; a multi-part launch seeds by
    the index within each part (a part's hits equal its own query); wavefront rounds, path tracing
    and the learner refuse it."""
    from nbvh_amd import hip
    cfg = nb.network_config(segment_distribution=nb.SEGMENT_UNIFORM)
    scene = nb.Scene(ctx, bench["v"], bench["t"])
    scene.build_nbvh(256)
    scene.init_network(cfg, seed=5, table_range=1e-1)
    cfg2, tables, mlp = scene.network()
    assert cfg2.segment_distribution == nb.SEGMENT_UNIFORM
    bmin, bmax = scene.bbox()
    net = orc.Net(cfg2, tables, mlp, bmin, bmax)
    mod = nb.NeuralBVHModule(ctx, scene)
    _, rays = camera_rays(nb, orc, 96, 96)
    rays = np.concatenate([rays, random_rays(5000, 21, inside_frac=0.4)])
    by_seed = {}
    for seed in (0, 4242):
        for any_hit in (False, True):
            kind = nb.QUERY_NEURAL_ANY if any_hit else nb.QUERY_NEURAL_CLOSEST
            got = mod.neural_bvh_traversal(rays, kind=kind, precision=nb.PREC_FP32, segment_seed=seed)
            exp = net.neural_query(scene.nbvh(), rays, any_hit=any_hit, mode=orc.FP32, segment_seed=seed)
            assert 0.1 < (nb.hit_kind(exp) == nb.HIT_NEURAL).mean() < 0.95
            assert hits_equal(got, exp), (seed, any_hit)
            by_seed[(seed, any_hit)] = got
        g16 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16, segment_seed=seed)
        e64 = net.neural_query(scene.nbvh(), rays, mode=orc.FP64, segment_seed=seed)
        assert (nb.hit_kind(g16) != nb.hit_kind(e64)).mean() <= 0.005
    assert not np.array_equal(by_seed[(0, False)]["t"], by_seed[(4242, False)]["t"])
    # two parts in one launch: each part's streams are seeded by the index within the part
    halves = [rays[:4000], rays[4000:]]
    bufs = [(hip.Buffer.from_array(r), hip.Buffer(len(r) * 48)) for r in halves]
    mod.query_parts_device(mod.params(precision=nb.PREC_FP32, segment_seed=9),
                           [(b.ptr, len(r), h.ptr, False) for (b, h), r in zip(bufs, halves)])
    hip.synchronize()
    for (b, h), r in zip(bufs, halves):
        assert hits_equal(h.download(nb.HIT_DTYPE)[:len(r)],
                          mod.neural_bvh_traversal(r, precision=nb.PREC_FP32, segment_seed=9))
    for call in (lambda: mod.neural_bvh_traversal(rays[:100], wavefront=True),
                 lambda: mod.path_trace(nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, 16, 16), 16, 16),
                 lambda: nb.Learner(ctx, scene)):
        with pytest.raises(nb.NbvhError) as e:
            call()
        assert e.value.status == nb._ffi.ERR_UNSUPPORTED
    scene.close()


def test_neural_query_step_cap_bit_exact(nb, orc, bench):
    rays = random_rays(4000, 8, inside_frac=0.5)
    got = bench["module"].neural_bvh_traversal(rays, max_traversal_steps=1)
    exp = bench["net"].neural_query(bench["nbvh"], rays, max_steps=1, mode=orc.FP32)
    assert hits_equal(got, exp)
    assert ((nb.hit_flags(got) & nb.FLAG_STEP_CAP) != 0).sum() > 0
    assert nb.hit_segments(got).max() == 2


def test_neural_query_threshold_bit_exact(nb, orc, bench):
    rays = random_rays(3000, 10, inside_frac=0.2)
    got = bench["module"].neural_bvh_traversal(rays, visibility_threshold=0.50001)
    exp = bench["net"].neural_query(bench["nbvh"], rays, threshold=0.50001, mode=orc.FP32)
    assert hits_equal(got, exp)


@pytest.mark.parametrize("lod", [0, 6])
def test_chess_nbvh_fp32_bit_exact(nb, orc, ctx, lod):
    cfg = nb.network_config(n_levels=8, n_features_per_level=4, log2_hashmap_size=18, base_resolution=8,
                            per_level_scale=2.0)
    scene = nb.Scene(ctx)
    scene.load_neural_bvh(CHESS)
    scene.init_network(cfg, seed=9, table_range=1e-4)
    _, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    nodes = scene.nbvh()
    assert len(nodes) == 25249
    net = orc.Net(cfg, tables, mlp, bmin, bmax)
    rng = np.random.default_rng(lod)
    n = 4000
    rays = np.zeros(n, orc.RAY_DTYPE)
    o = rng.uniform(bmin - 0.2, bmax + 0.2, (n, 3))
    tgt = rng.uniform(bmin, bmax, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"], rays["d"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, np.float32(3.4028234663852886e38)
    got = nb.NeuralBVHModule(ctx, scene).neural_bvh_traversal(rays, lod=lod)
    exp = net.neural_query(nodes, rays, lod=lod, mode=orc.FP32)
    assert hits_equal(got, exp)
    assert nb.hit_segments(got).mean() > 0.5
    scene.close()


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
def test_network_mfma_vs_fp64(nb, orc, bench, precision):
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}[precision]
    # Segment inputs of real queries (the oracle's segment log), encoded in FP32.
    _, rays = camera_rays(nb, orc, 96, 96)
    _, log = bench["net"].neural_query(bench["nbvh"], rays, mode=orc.FP64, log_segments=True)
    x = bench["net"].encode(rays[log["ray"]], log["mint"], log["maxt"], mode=orc.FP32)
    out, logits = bench["module"].network_inference(x, precision=prec)
    o64, z64 = bench["net"].network(x, mode=orc.FP64)
    scale = np.abs(z64).max(axis=0)
    err = np.abs(logits - z64) / scale
    # Relative to each output's range: FP16 keeps ~11 bits, BF16 ~8 bits through 5 layers, FP8
    # e4m3 ~4 bits through its 3 hidden layers (measured max 0.14, flips 2.2-2.4 %).
    tol = {"fp16": 2e-2, "bf16": 8e-2, "fp8": 0.25}[precision]
    assert err.max() < tol, err.max(axis=0)
    flips = np.mean(np.sign(logits[:, 0]) != np.sign(z64[:, 0]))
    assert flips <= {"fp16": 0.005, "bf16": 0.005, "fp8": 0.05}[precision], flips
    assert len(x) > 5000


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
def test_neural_query_vs_fp64(nb, orc, bench, precision):
    # Bars: hit/miss disagreement with the FP64 oracle <= 0.5 % (FP16, BF16; the headline bar) or
    # <= 5 % (FP8 e4m3 hidden layers); p99 relative depth error on agreeing hits.
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}[precision]
    _, rays = camera_rays(nb, orc, 128, 128)
    rays = np.concatenate([rays, random_rays(4000, 3, inside_frac=0.3)])
    got = bench["module"].neural_bvh_traversal(rays, precision=prec)
    exp = bench["net"].neural_query(bench["nbvh"], rays, mode=orc.FP64)
    kg, ke = nb.hit_kind(got), nb.hit_kind(exp)
    assert (kg != ke).mean() <= {"fp16": 0.005, "bf16": 0.005, "fp8": 0.05}[precision], (kg != ke).mean()
    same = (kg == 1) & (ke == 1) & (got["node_id"] == exp["node_id"])
    assert same.mean() > 0.3
    rel = np.abs(got["t"][same] - exp["t"][same]) / np.maximum(exp["t"][same], 1e-3)
    assert np.percentile(rel, 99) < {"fp16": 1e-2, "bf16": 3e-2, "fp8": 0.1}[precision], np.percentile(rel, 99)


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
def test_bounce_rays_vs_fp64(nb, orc, bench, precision):
    """Diffuse bounce rays (the path tracer's secondary rays: origins on the surface, inside a
    leaf box, cosine-distributed directions) against the FP64 oracle, with the same bars as the
    primary rays: hit/miss disagreement <= 0.5 % (FP16, BF16) or <= 5 % (FP8), p99 relative depth
    error on agreeing hits."""
    from nbvh_amd import hip
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}[precision]
    _, rays = camera_rays(nb, orc, 160, 120)
    hits = bench["module"].neural_bvh_traversal(rays, precision=nb.PREC_FP32)
    dr, dh, dn = hip.Buffer.from_array(rays), hip.Buffer.from_array(hits), hip.Buffer(len(rays) * 32)
    bench["module"].scatter_diffuse_device(dr.ptr, dh.ptr, len(rays), dn.ptr, sample=3, bounce=1)
    hip.synchronize()
    bounce = dn.download(nb.RAY_DTYPE)
    bounce = bounce[bounce["tmax"] > 0]
    assert len(bounce) > 5000
    got = bench["module"].neural_bvh_traversal(bounce, precision=prec)
    exp = bench["net"].neural_query(bench["nbvh"], bounce, mode=orc.FP64)
    kg, ke = nb.hit_kind(got), nb.hit_kind(exp)
    assert (kg != ke).mean() <= {"fp16": 0.005, "bf16": 0.005, "fp8": 0.05}[precision], (kg != ke).mean()
    same = (kg == 1) & (ke == 1) & (got["node_id"] == exp["node_id"])
    assert same.sum() > 500
    rel = np.abs(got["t"][same] - exp["t"][same]) / np.maximum(exp["t"][same], 1e-3)
    assert np.percentile(rel, 99) < {"fp16": 1e-2, "bf16": 3e-2, "fp8": 0.1}[precision], np.percentile(rel, 99)


def test_errors_through_abi(nb, ctx):
    scene = nb.Scene(ctx)
    mod = nb.NeuralBVHModule(ctx, scene)
    rays = random_rays(10, 1)
    with pytest.raises(nb.NbvhError) as e:
        mod.neural_bvh_traversal(rays)
    assert e.value.status == 6  # no N-BVH
    with pytest.raises(nb.NbvhError) as e:
        mod.neural_bvh_traversal(rays, kind=nb.QUERY_TRIANGLE_CLOSEST)
    assert e.value.status == 6  # no triangles
    scene.set_nbvh(nb.nbvh_cut(nb.bvh_build(*nb.icosphere(2))[0], 8))
    with pytest.raises(nb.NbvhError) as e:
        mod.neural_bvh_traversal(rays)
    assert e.value.status == 6  # no weights
    scene.init_network(nb.network_config(n_levels=2, log2_hashmap_size=10), seed=1)
    mod.neural_bvh_traversal(rays)  # FP32 kernel exists for this shape
    with pytest.raises(nb.NbvhError) as e:
        mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16)
    assert e.value.status == 7  # no MFMA kernel for K0 = 12
    scene.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_cpp_driver_matches_python_path(nb, orc, bench, tmp_path, precision):
    """The C++ host driver (include/nbvh.hpp -> C ABI) produces the same hit records as the
    Python host path for the same scene, camera and precision."""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "nbvh_amd", "nbvh_driver")
    out = tmp_path / "hits.bin"
    r = subprocess.run([exe, "query", "--width", "256", "--height", "192", "--precision", precision, "--out",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, nb.HIT_DTYPE)
    _, rays = camera_rays(nb, orc, 256, 192)
    prec = {"fp32": nb.PREC_FP32, "fp16": nb.PREC_FP16}[precision]
    exp = bench["module"].neural_bvh_traversal(rays, precision=prec)
    assert hits_equal(got, exp)
    assert (nb.hit_kind(got) == nb.HIT_NEURAL).mean() > 0.2


@pytest.mark.parametrize("family,precision", [("neural", "fp16"), ("neural", "fp32"), ("triangle", "fp32")])
def test_query_parts_mixed_kinds(nb, orc, bench, family, precision):
    """One launch over three ray arrays (closest, any, closest, with an empty part between):
    each part's hits equal a direct query of its own kind."""
    from nbvh_amd import hip
    prec = {"fp16": nb.PREC_FP16, "fp32": nb.PREC_FP32}[precision]
    closest = nb.QUERY_NEURAL_CLOSEST if family == "neural" else nb.QUERY_TRIANGLE_CLOSEST
    any_k = nb.QUERY_NEURAL_ANY if family == "neural" else nb.QUERY_TRIANGLE_ANY
    arrays = [random_rays(3000, 90, inside_frac=0.3), random_rays(2500, 91, inside_frac=0.5),
              random_rays(0, 92), random_rays(1700, 93)]
    kinds = [False, True, True, False]
    bufs = [(hip.Buffer.from_array(r) if len(r) else None, hip.Buffer(max(1, len(r)) * 48)) for r in arrays]
    parts = [((b.ptr if b else 0), len(r), h.ptr, k) for (b, h), r, k in zip(bufs, arrays, kinds)]
    mod = bench["module"]
    mod.query_parts_device(mod.params(kind=closest, precision=prec), parts)
    hip.synchronize()
    for (b, h), r, k in zip(bufs, arrays, kinds):
        if not len(r):
            continue
        got = h.download(nb.HIT_DTYPE)[:len(r)]
        exp = mod.neural_bvh_traversal(r, kind=any_k if k else closest, precision=prec)
        assert hits_equal(got, exp)


def test_graph_replays_match_eager(nb, orc, bench):
    """A multi-part query captured into a HIP graph (counter reset kernel + part table copy +
    query kernel) gives the eager result on every replay, with eager work in between."""
    from nbvh_amd import hip
    mod = bench["module"]
    _, cam = camera_rays(nb, orc, 128, 96)
    shadow = random_rays(4000, 95, inside_frac=0.4)
    dc, ds = hip.Buffer.from_array(cam), hip.Buffer.from_array(shadow)
    hc, hs = hip.Buffer(len(cam) * 48), hip.Buffer(len(shadow) * 48)
    p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16)
    parts = [(dc.ptr, len(cam), hc.ptr, False), (ds.ptr, len(shadow), hs.ptr, True)]
    s = hip.Stream()
    mod.query_parts_device(p, parts, stream=s.handle)  # eager (also allocates the part tables)
    s.synchronize()
    ref_c, ref_s = hc.download(nb.HIT_DTYPE), hs.download(nb.HIT_DTYPE)
    with hip.Graph.capture(s) as g:
        mod.query_parts_device(p, parts, stream=s.handle)
    for rep in range(4):
        hc.zero()
        hs.zero()
        g.launch(s)
        s.synchronize()
        assert hits_equal(hc.download(nb.HIT_DTYPE), ref_c), rep
        assert hits_equal(hs.download(nb.HIT_DTYPE), ref_s), rep
        mod.neural_bvh_traversal(shadow[:500], precision=nb.PREC_FP32)  # eager work between replays
    g.destroy()
    s.destroy()


def test_scratch_slots_are_not_shared(nb, orc, bench):
    """Per-query scratch (queue counters and part tables): a multi-part query recorded into a
    graph keeps its own slot, so replays stay equal to the eager result after 2,100 further
    multi-part queries (more than the 2,048 round-robin slots, each with its own part table);
    those 2,100 queries, queued on 4 streams without waiting, each give the eager result of
    their own rays although far more than one ring of them is in flight."""
    from nbvh_amd import hip
    mod = bench["module"]
    _, cam = camera_rays(nb, orc, 96, 64)
    shadow = random_rays(len(cam), 97, inside_frac=0.4)
    n_q, m = 2100, 64
    # every window [8 (i % 700), 8 (i % 700) + m) of the small queries lies inside both ray arrays
    assert 8 * 699 + m <= min(len(cam), len(shadow))
    p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16)
    dc, ds = hip.Buffer.from_array(cam), hip.Buffer.from_array(shadow)
    hc, hs = hip.Buffer(len(cam) * 48), hip.Buffer(len(shadow) * 48)
    parts = [(dc.ptr, len(cam), hc.ptr, False), (ds.ptr, len(shadow), hs.ptr, True)]
    s = hip.Stream()
    mod.query_parts_device(p, parts, stream=s.handle)
    s.synchronize()
    ref_c, ref_s = hc.download(nb.HIT_DTYPE), hs.download(nb.HIT_DTYPE)
    with hip.Graph.capture(s) as g:
        mod.query_parts_device(p, parts, stream=s.handle)
    # 2,100 small two-part queries: rays [8i, 8i + 64) of the camera rays as closest hits and of the
    # shadow rays as any hits, each into its own hit arrays.
    out_c, out_s = hip.Buffer(n_q * m * 48), hip.Buffer(n_q * m * 48)
    streams = [hip.Stream() for _ in range(4)]
    for i in range(n_q):
        o = 8 * (i % 700)
        mod.query_parts_device(p, [(dc.ptr + 32 * o, m, out_c.ptr + 48 * m * i, False),
                                   (ds.ptr + 32 * o, m, out_s.ptr + 48 * m * i, True)],
                               stream=streams[i % 4].handle)
    hip.synchronize()
    got_c = out_c.download(nb.HIT_DTYPE).reshape(n_q, m)
    got_s = out_s.download(nb.HIT_DTYPE).reshape(n_q, m)
    for i in range(n_q):
        o = 8 * (i % 700)
        assert hits_equal(got_c[i], ref_c[o:o + m]) and hits_equal(got_s[i], ref_s[o:o + m]), i
    for rep in range(3):
        hc.zero()
        hs.zero()
        g.launch(s)
        s.synchronize()
        assert hits_equal(hc.download(nb.HIT_DTYPE), ref_c), rep
        assert hits_equal(hs.download(nb.HIT_DTYPE), ref_s), rep
    g.destroy()
    for st in streams:
        st.destroy()
    s.destroy()


def test_captured_query_slots_are_bounded(nb, orc):
    """A context hands at most 1,024 scratch slots to queries recorded into graphs (they are never
    reused); the next capture fails with NBVH_ERR_CAPACITY before it enqueues anything, and eager
    queries keep working."""
    from nbvh_amd import hip
    ctx = nb.Context()
    v, t = nb.icosphere(2, seed=7, amplitude=0.1)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(16)
    scene.init_network(nb.network_config(log2_hashmap_size=14), seed=1, table_range=1e-4)
    mod = nb.NeuralBVHModule(ctx, scene)
    rays = random_rays(256, 5, inside_frac=0.2)
    dr, dh = hip.Buffer.from_array(rays), hip.Buffer(len(rays) * 48)
    p = mod.params(precision=nb.PREC_FP16)
    s = hip.Stream()
    with hip.Graph.capture(s) as g:
        for _ in range(1024):
            mod.query_device(p, dr.ptr, len(rays), dh.ptr, stream=s.handle)
        with pytest.raises(nb.NbvhError) as e:
            mod.query_device(p, dr.ptr, len(rays), dh.ptr, stream=s.handle)
        assert e.value.status == nb._ffi.ERR_CAPACITY
    g.launch(s)
    s.synchronize()
    exp = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16)
    assert hits_equal(dh.download(nb.HIT_DTYPE), exp)
    g.destroy()
    s.destroy()
    scene.close()
    ctx.close()


@pytest.mark.parametrize("kind", ["neural", "triangle"])
def test_scatter_diffuse_matches_oracle(nb, orc, bench, kind):
    """The bounce kernel against its restatement: same inactive rays, origins within 2e-6 and
    directions within 1e-5. Device and host sin/cos differ by a few ulp, and near the horizon the
    hemisphere's z = sqrt(1 - x^2 - y^2) amplifies that (1 - r^2 cancels): 1e-5 is 6e-4 degrees."""
    from nbvh_amd import hip
    _, rays = camera_rays(nb, orc, 160, 120)
    k = nb.QUERY_NEURAL_CLOSEST if kind == "neural" else nb.QUERY_TRIANGLE_CLOSEST
    hits = bench["module"].neural_bvh_traversal(rays, kind=k, precision=nb.PREC_FP32)
    dr, dh, dn = hip.Buffer.from_array(rays), hip.Buffer.from_array(hits), hip.Buffer(len(rays) * 32)
    bench["module"].scatter_diffuse_device(dr.ptr, dh.ptr, len(rays), dn.ptr, sample=2, bounce=1)
    hip.synchronize()
    got = dn.download(nb.RAY_DTYPE)
    exp = orc.scatter_diffuse(rays, hits, sample=2, bounce=1, verts=bench["v"], tris=bench["t"])
    assert np.array_equal(got["tmax"], exp["tmax"])
    live = exp["tmax"] > 0
    assert live.mean() > 0.2
    assert np.abs(got["d"][live] - exp["d"][live]).max() < 1e-5
    assert np.abs(got["o"][live] - exp["o"][live]).max() < 2e-6


def test_launch_shapes_and_autotune_keep_results(nb, orc, bench):
    """Every launch shape (workgroups per CU x queue bands x ray order) gives identical hits; the
    autotuner keeps one of the candidates; a 65,536-ray batch without an image order is sorted
    where the shape says so, and NBVH_QUERY_ARRAY_ORDER keeps array order either way."""
    from nbvh_amd import hip
    mod = bench["module"]
    _, rays = camera_rays(nb, orc, 256, 256)
    rng = np.random.default_rng(4)
    rays = rays[rng.permutation(len(rays))]  # incoherent order
    dr, dh = hip.Buffer.from_array(rays), hip.Buffer(len(rays) * 48)
    ref = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16, array_order=True)
    saved = mod.launch_shape(nb.PREC_FP16)
    for bpc, bands, order in [(1, 1, 0), (2, 8, 1), (1, 4, 0), (3, 2, 1), (0, 8, 1)]:
        mod.set_launch_shape(nb.PREC_FP16, bpc, bands, order)
        assert mod.launch_shape(nb.PREC_FP16) == (bpc, bands, order)
        assert hits_equal(mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16), ref)
        assert hits_equal(mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16, array_order=True), ref)
    (bpc, bands, order), ms = mod.autotune(mod.params(precision=nb.PREC_FP16), dr.ptr, len(rays), dh.ptr)
    assert bpc in (0, 1, 2, 3) and bands in (1, 2, 4, 8) and order in (0, 1) and ms > 0
    assert mod.launch_shape(nb.PREC_FP16) == (bpc, bands, order)
    assert hits_equal(dh.download(nb.HIT_DTYPE), ref)
    with pytest.raises(nb.NbvhError):
        mod.set_launch_shape(nb.PREC_FP16, 0, 0, 2)
    mod.set_launch_shape(nb.PREC_FP16, *saved)


def test_segment_labels_match_oracle(nb, orc, bench):
    """GPU label kernel vs the restatement on real segments: visibility and depth bit-exact,
    normals within 1e-7."""
    rays = np.concatenate([camera_rays(nb, orc, 96, 96)[1], random_rays(3000, 97, inside_frac=0.3)])
    _, log = bench["net"].neural_query(bench["nbvh"], rays, threshold=0.999999, mode=orc.FP32, log_segments=True)
    r = rays[log["ray"]]
    mint = np.maximum(log["mint"], 0).astype(np.float32)
    got = bench["module"].segment_labels(r, mint, log["maxt"])
    exp = orc.segment_labels(bench["v"], bench["t"], bench["nodes"], bench["prims"], r, mint, log["maxt"])
    assert (exp[:, 0] == 1).mean() > 0.1 and len(r) > 5000
    assert np.array_equal(got[:, :2], exp[:, :2])
    assert np.abs(got[:, 2:] - exp[:, 2:]).max() <= 1e-7


@pytest.mark.parametrize("kind,precision", [("neural", "fp16"), ("neural", "fp32"), ("triangle", "fp32")])
def test_sorted_ray_order_same_results(nb, orc, bench, kind, precision):
    """NBVH_QUERY_SORT_RAYS only changes the tracing order: hits are identical, for one array and
    for a multi-part launch (each part sorted within itself)."""
    from nbvh_amd import hip
    mod = bench["module"]
    prec = {"fp16": nb.PREC_FP16, "fp32": nb.PREC_FP32}[precision]
    k = nb.QUERY_NEURAL_CLOSEST if kind == "neural" else nb.QUERY_TRIANGLE_CLOSEST
    rays = random_rays(20000, 98, inside_frac=0.3)
    a = mod.neural_bvh_traversal(rays, kind=k, precision=prec)
    b = mod.neural_bvh_traversal(rays, kind=k, precision=prec, sort_rays=True)
    assert hits_equal(a, b)
    r2 = random_rays(7000, 99, inside_frac=0.5)
    d1, d2 = hip.Buffer.from_array(rays), hip.Buffer.from_array(r2)
    h1, h2 = hip.Buffer(len(rays) * 48), hip.Buffer(len(r2) * 48)
    mod.query_parts_device(mod.params(kind=k, precision=prec, sort_rays=True),
                           [(d1.ptr, len(rays), h1.ptr, False), (d2.ptr, len(r2), h2.ptr, True)])
    hip.synchronize()
    assert hits_equal(h1.download(nb.HIT_DTYPE), a)
    any_k = nb.QUERY_NEURAL_ANY if kind == "neural" else nb.QUERY_TRIANGLE_ANY
    assert hits_equal(h2.download(nb.HIT_DTYPE), mod.neural_bvh_traversal(r2, kind=any_k, precision=prec))


def hybrid_environment(nb):
    """Triangle meshes around the neural icosphere: a floor quad and four geodesic spheres."""
    vs, ts, off = [], [], 0
    quad_v = np.array([[-3, -1.3, -3], [3, -1.3, -3], [3, -1.3, 3], [-3, -1.3, 3]], np.float32)
    quad_t = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    vs.append(quad_v)
    ts.append(quad_t)
    off = 4
    for i, c in enumerate([(1.6, -0.8, 0.2), (-1.5, -0.7, -0.4), (0.3, 1.5, -0.9), (-0.2, -0.9, 1.6)]):
        v, t = nb.geosphere(12, seed=40 + i, amplitude=0.05, center=c, radius=0.45)
        vs.append(v)
        ts.append(t + np.uint32(off))
        off += len(v)
    return np.concatenate(vs), np.concatenate(ts)


@pytest.mark.parametrize("any_hit", [False, True])
def test_hybrid_query_matches_oracle(nb, orc, ctx, bench, any_hit):
    """Hybrid closest/any hit over the neural icosphere and a triangle environment: FP32 bit-exact
    with the oracle's composition; the hybrid t never exceeds the triangle t; both kinds occur."""
    ev, et = hybrid_environment(nb)
    env = nb.Scene(ctx, ev, et)
    enodes, eprims = env.bvh()
    _, cam = camera_rays(nb, orc, 128, 128, eye=(0.5, 0.6, 3.2))
    rays = np.concatenate([cam, random_rays(6000, 101, origin_box=3.0, inside_frac=0.2)])
    kind = nb.QUERY_NEURAL_ANY if any_hit else nb.QUERY_NEURAL_CLOSEST
    got = bench["module"].query_hybrid(env, rays, kind=kind, precision=nb.PREC_FP32)
    exp = orc.hybrid_query(bench["net"], bench["nbvh"], ev, et, enodes, eprims, rays, any_hit=any_hit)
    assert hits_equal(got, exp)
    kinds = nb.hit_kind(got)
    assert (kinds == nb.HIT_NEURAL).mean() > 0.05 and (kinds == nb.HIT_TRIANGLE).mean() > 0.05
    if not any_hit:
        tenv = orc.tri_query(ev, et, enodes, eprims, rays)
        assert np.all(got["t"] <= tenv["t"])
    g16 = bench["module"].query_hybrid(env, rays, kind=kind, precision=nb.PREC_FP16)
    e64 = orc.hybrid_query(bench["net"], bench["nbvh"], ev, et, enodes, eprims, rays, any_hit=any_hit,
                           mode=orc.FP64)
    assert (nb.hit_kind(g16) != nb.hit_kind(e64)).mean() <= 0.005
    env.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "fp8"])
def test_wavefront_mode_same_results(nb, orc, bench, precision):
    """Sorted wavefront rounds (NBVH_QUERY_WAVEFRONT) give the persistent kernel's hit records
    bit for bit: closest and any hit, step cap, threshold, and a multi-part launch."""
    from nbvh_amd import hip
    mod = bench["module"]
    prec = {"fp32": nb.PREC_FP32, "fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}[precision]
    _, cam = camera_rays(nb, orc, 96, 64)
    rays = np.concatenate([cam, random_rays(9000, 103, inside_frac=0.4)])
    for kw in ({}, {"kind": nb.QUERY_NEURAL_ANY}, {"max_traversal_steps": 1}, {"visibility_threshold": 0.50001}):
        a = mod.neural_bvh_traversal(rays, precision=prec, **kw)
        b = mod.neural_bvh_traversal(rays, precision=prec, wavefront=True, **kw)
        assert hits_equal(a, b), kw
    if precision == "fp32":
        assert hits_equal(b, bench["net"].neural_query(bench["nbvh"], rays, threshold=0.50001, mode=orc.FP32))
    r2 = random_rays(3000, 104, inside_frac=0.5)
    d1, d2 = hip.Buffer.from_array(rays), hip.Buffer.from_array(r2)
    h1, h2 = hip.Buffer(len(rays) * 48), hip.Buffer(len(r2) * 48)
    mod.query_parts_device(mod.params(precision=prec, wavefront=True),
                           [(d1.ptr, len(rays), h1.ptr, False), (d2.ptr, len(r2), h2.ptr, True)])
    hip.synchronize()
    assert hits_equal(h1.download(nb.HIT_DTYPE), mod.neural_bvh_traversal(rays, precision=prec))
    assert hits_equal(h2.download(nb.HIT_DTYPE),
                      mod.neural_bvh_traversal(r2, kind=nb.QUERY_NEURAL_ANY, precision=prec))


def edge_rays(seed):
    """Degenerate and boundary rays: axis-aligned directions (two zero components), one zero
    component, origins exactly on the scene box faces, short tmax (stops before, inside or just
    past the surface), empty intervals (tmin > tmax) and origins inside the object."""
    from oracle import RAY_DTYPE
    rng = np.random.default_rng(seed)
    out = []
    axes = np.concatenate([np.eye(3), -np.eye(3)])
    for k in range(6):  # axis-aligned, origins off-axis in the box
        o = rng.uniform(-1.2, 1.2, (300, 3))
        o[:, k % 3] = 1.3 * (1 if k < 3 else -1) * -1
        d = np.tile(axes[k], (300, 1))
        out.append((o, d, 1e-4, np.float32(3.4028234663852886e38)))
    o = rng.uniform(-1.5, 1.5, (600, 3))
    d = rng.normal(size=(600, 3))
    d[np.arange(600), rng.integers(0, 3, 600)] = 0.0  # one zero component
    out.append((o, d, 1e-4, np.float32(3.4028234663852886e38)))
    o = rng.uniform(-1.5, 1.5, (600, 3))
    d = -o + rng.normal(scale=0.2, size=(600, 3))
    out.append((o, d, 1e-4, rng.uniform(0.05, 2.5, 600)))  # short rays
    out.append((o[:200], d[:200], 1.0, 0.5))  # empty interval
    o = rng.uniform(-0.3, 0.3, (400, 3))  # inside the object
    out.append((o, rng.normal(size=(400, 3)), 1e-4, np.float32(3.4028234663852886e38)))
    rays = []
    for o, d, tmin, tmax in out:
        r = np.zeros(len(o), RAY_DTYPE)
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        r["o"], r["d"] = o.astype(np.float32), d.astype(np.float32)
        r["tmin"], r["tmax"] = tmin, tmax
        rays.append(r)
    return np.concatenate(rays)


def test_edge_rays_bit_exact(nb, orc, bench):
    rays = edge_rays(31)
    # Origins exactly on the scene box faces, moving along the face plane and into it.
    bmin, bmax = bench["scene"].bbox()
    face = np.zeros(300, orc.RAY_DTYPE)
    rng = np.random.default_rng(3)
    o = rng.uniform(bmin, bmax, (300, 3))
    o[:150, 0], o[150:, 1] = bmin[0], bmax[1]
    d = np.zeros((300, 3))
    d[:150, 0], d[150:, 1] = 1.0, -1.0
    d[::2, 2] = 0.3
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    face["o"], face["d"] = o.astype(np.float32), d.astype(np.float32)
    face["tmin"], face["tmax"] = 0.0, np.float32(3.4028234663852886e38)
    rays = np.concatenate([rays, face])
    mod = bench["module"]
    for any_hit in (False, True):
        kind = nb.QUERY_NEURAL_ANY if any_hit else nb.QUERY_NEURAL_CLOSEST
        got = mod.neural_bvh_traversal(rays, kind=kind, precision=nb.PREC_FP32)
        exp = bench["net"].neural_query(bench["nbvh"], rays, any_hit=any_hit, mode=orc.FP32)
        assert hits_equal(got, exp)
        tkind = nb.QUERY_TRIANGLE_ANY if any_hit else nb.QUERY_TRIANGLE_CLOSEST
        tgot = mod.neural_bvh_traversal(rays, kind=tkind)
        texp = orc.tri_query(bench["v"], bench["t"], bench["nodes"], bench["prims"], rays, any_hit=any_hit)
        assert np.array_equal(tgot["info"], texp["info"])
        if not any_hit:
            assert hits_equal(tgot, texp)
    # Empty intervals never hit; the short rays end both ways.
    k = nb.hit_kind(got)
    assert 0.05 < (k != 0).mean() < 0.95


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp8"])
def test_batch_size_independence(nb, orc, bench, precision):
    """A ray's record does not depend on the batch around it: the first n rays of a batch, traced
    alone for empty, single, ragged and wave-multiple n, give the records of the full batch."""
    prec = {"fp32": nb.PREC_FP32, "fp16": nb.PREC_FP16, "fp8": nb.PREC_FP8}[precision]
    _, cam = camera_rays(nb, orc, 64, 48)
    rays = np.concatenate([random_rays(1500, 12, inside_frac=0.3), cam])
    mod = bench["module"]
    full = mod.neural_bvh_traversal(rays, precision=prec)
    for n in (0, 1, 2, 63, 64, 65, 255, 257, 1000):
        got = mod.neural_bvh_traversal(rays[:n], precision=prec)
        assert len(got) == n and hits_equal(got, full[:n]), n
    rev = mod.neural_bvh_traversal(rays[::-1].copy(), precision=prec)
    assert hits_equal(rev[::-1], full)


def test_tcnn_model_file_loads_the_same_network(nb, orc, ctx, bench, tmp_path):
    """A model written in the reference's serialised form (tiny-cuda-nn flat order, JSON text)
    and loaded with Scene.load_tcnn_model traces the same records as the original weights."""
    import json
    cfg, tables, mlp = bench["scene"].network()
    params = nb.network_to_tcnn_params(cfg, tables, mlp)
    path = tmp_path / "model.json"
    path.write_text(json.dumps({"n_params": len(params), "params_binary": {"bytes": list(params.tobytes())}}))
    v, t = bench["v"], bench["t"]
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.load_tcnn_model(cfg, str(path))
    _, rays = camera_rays(nb, orc, 64, 64)
    for prec in (nb.PREC_FP32, nb.PREC_FP16):
        a = bench["module"].neural_bvh_traversal(rays, precision=prec)
        b = nb.NeuralBVHModule(ctx, scene).neural_bvh_traversal(rays, precision=prec)
        assert hits_equal(a, b)
    scene.close()


@pytest.fixture(scope="module")
def chess(nb, orc, ctx):
    """The reference's chess N-BVH (25,249 nodes) with a seeded network of the shipped chess
    configuration: 8 levels x 4 features, 2^18 entries, base 8, scale 2 (the (3, 8, 4) MFMA
    kernels, 8-byte corner loads)."""
    cfg = nb.network_config(n_levels=8, n_features_per_level=4, log2_hashmap_size=18, base_resolution=8,
                            per_level_scale=2.0)
    scene = nb.Scene(ctx)
    scene.load_neural_bvh(CHESS)
    scene.init_network(cfg, seed=9, table_range=1e-2)
    _, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    rng = np.random.default_rng(17)
    n = 6000
    rays = np.zeros(n, orc.RAY_DTYPE)
    o = rng.uniform(bmin - 0.2, bmax + 0.2, (n, 3))
    d = rng.uniform(bmin, bmax, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["o"], rays["d"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, np.float32(3.4028234663852886e38)
    yield dict(scene=scene, net=orc.Net(cfg, tables, mlp, bmin, bmax), nodes=scene.nbvh(), rays=rays,
               mod=nb.NeuralBVHModule(ctx, scene))
    scene.close()


@pytest.mark.parametrize("precision,lod", [("fp16", 0), ("fp16", 6), ("bf16", 0)])
def test_chess_mfma_query_vs_fp64(nb, orc, chess, precision, lod):
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16}[precision]
    got = chess["mod"].neural_bvh_traversal(chess["rays"], precision=prec, lod=lod)
    exp = chess["net"].neural_query(chess["nodes"], chess["rays"], lod=lod, mode=orc.FP64)
    kg, ke = nb.hit_kind(got), nb.hit_kind(exp)
    assert 0.05 < (ke == 1).mean() < 0.95
    assert (kg != ke).mean() <= 0.005
    same = (kg == 1) & (ke == 1) & (got["node_id"] == exp["node_id"])
    rel = np.abs(got["t"][same] - exp["t"][same]) / np.maximum(exp["t"][same], 1e-3)
    assert np.percentile(rel, 99) < 2e-2


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
def test_chess_network_mfma_vs_fp64(nb, orc, chess, precision):
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}[precision]
    _, log = chess["net"].neural_query(chess["nodes"], chess["rays"], mode=orc.FP64, log_segments=True)
    x = chess["net"].encode(chess["rays"][log["ray"]], log["mint"], log["maxt"], mode=orc.FP32)
    assert len(x) > 4000
    _, z = chess["mod"].network_inference(x, precision=prec)
    _, z64 = chess["net"].network(x, mode=orc.FP64)
    err = np.abs(z - z64) / np.abs(z64).max(axis=0)
    assert err.max() < {"fp16": 2e-2, "bf16": 8e-2, "fp8": 0.25}[precision], err.max(axis=0)
    flips = np.mean(np.sign(z[:, 0]) != np.sign(z64[:, 0]))
    assert flips <= {"fp16": 0.005, "bf16": 0.005, "fp8": 0.05}[precision], flips


def multi_lod_nbvh(nb, nodes, n_leaves=2048, marks=((3, 6), (5, 3))):
    """An N-BVH cut whose interior nodes at the given depths are also LoD leaves, children kept
    (the form the reference's LoD assignment leaves): marks = ((depth, lod), ...)."""
    nbvh = nb.nbvh_cut(nodes, n_leaves)
    depth = np.full(len(nbvh), -1)
    depth[0] = 0
    for i in range(len(nbvh)):
        if nbvh["prim_count"][i] == 0 and nbvh["left_or_prim_offset"][i] != nb.NO_NEURAL_CHILD:
            c = int(nbvh["left_or_prim_offset"][i])
            depth[c] = depth[c + 1] = depth[i] + 1
            for d, lod in marks:
                if depth[i] == d:
                    nbvh["prim_count"][i] = 2 + lod
    assert np.all(depth >= 0)
    return nbvh


@pytest.mark.parametrize("lod", [0, 3, 6])
def test_lod_selection_parity(nb, orc, ctx, bench, lod):
    """Traversal at LoD l stops at interior nodes marked LEAF_LOD(<= l): FP32 bit-exact with the
    oracle and FP16 within the visibility bar, on an N-BVH with LoD-marked interior nodes at two
    depths; the leaf sets differ between LoDs."""
    scene = nb.Scene(ctx, bench["v"], bench["t"])
    scene.set_nbvh(multi_lod_nbvh(nb, bench["nodes"]))
    scene.init_network(nb.network_config(), seed=4, table_range=1e-3)
    cfg, tables, mlp = scene.network()
    lo, hi = scene.bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    nodes = scene.nbvh()
    _, cam = camera_rays(nb, orc, 96, 96)
    rays = np.concatenate([cam, random_rays(3000, 71, inside_frac=0.3)])
    mod = nb.NeuralBVHModule(ctx, scene)
    got = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP32, lod=lod)
    assert hits_equal(got, net.neural_query(nodes, rays, lod=lod, mode=orc.FP32))
    g16 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16, lod=lod)
    e64 = net.neural_query(nodes, rays, lod=lod, mode=orc.FP64)
    assert (nb.hit_kind(g16) != nb.hit_kind(e64)).mean() <= 0.005
    if lod > 0:
        g0 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP32, lod=0)
        assert not np.array_equal(g0["node_id"], got["node_id"])
    scene.close()
