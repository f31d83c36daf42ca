"""GPU path-trace driver (nbvh_path_trace) against the oracle's path-trace restatement.

The FP32 driver traces the same paths as the oracle: primary hits and the first bounce's rays are
identical (the FP32 query is bit-exact); bounce directions come from the device's cosf/sinf, which
can differ from the host's in the last bit, so a path that grazes an edge may take another way
later. Bars: every pixel written exactly once per sample; rays per depth equal at depths 1-2 and
within 0.5 % after; >= 97 % of the pixels within 1e-5 of the oracle and the mean within 2e-3.
"""
import numpy as np
import pytest

from conftest import bench_scene

pytestmark = pytest.mark.gpu

W, H = 96, 64


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(nb, orc, ctx):
    v, t, nodes, prims, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(nb.network_config(), seed=1, table_range=1e-4)
    cfg, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    net = orc.Net(cfg, tables, mlp, bmin, bmax)
    import bench
    ev, et = bench.hybrid_environment(nb)
    env = nb.Scene(ctx, ev, et)
    enodes, eprims = env.bvh()
    yield dict(scene=scene, net=net, nbvh=scene.nbvh(), mod=nb.NeuralBVHModule(ctx, scene), env=env, ev=ev, et=et,
               enodes=enodes, eprims=eprims, v=v, t=t, nodes=nodes, prims=prims)
    env.close()
    scene.close()


def compare(gpu, ref, gpu_counts, ref_counts):
    g, r = gpu.reshape(-1, 4), ref.reshape(-1, 4)
    assert np.array_equal(g[:, 3], r[:, 3])
    assert gpu_counts[:2] == ref_counts[:2]
    for a, b in zip(gpu_counts[2:], ref_counts[2:]):
        assert abs(a - b) <= max(2, 0.005 * b)
    close = np.all(np.abs(g[:, :3] - r[:, :3]) <= 1e-5, axis=1)
    assert close.mean() >= 0.97, close.mean()
    assert abs(float(g[:, :3].mean() - r[:, :3].mean())) < 2e-3


@pytest.mark.parametrize("rr", [0, 3])
def test_path_trace_fp32_matches_oracle(nb, orc, setup, rr):
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    img, counts = setup["mod"].path_trace(cam, W, H, precision=nb.PREC_FP32, max_depth=5, rr_start_depth=rr)
    q = lambda r: setup["net"].neural_query(setup["nbvh"], r, mode=orc.FP32)  # noqa: E731
    ref, rc = orc.path_trace(q, nb.camera_array(cam), W, H, max_depth=5, rr_start_depth=rr)
    compare(img, ref, counts[0], rc)
    assert (img[..., 3] == 1).all() and img[..., :3].max() > 0.1


def test_hybrid_path_trace_fp32_matches_oracle(nb, orc, setup):
    cam = nb.camera_look_at((0.5, 0.6, 3.2), (0, -0.2, 0), (0, 1, 0), 45.0, W, H)
    albedo = (0.7, 0.6, 0.5)
    img, counts = setup["mod"].path_trace(cam, W, H, environment=setup["env"], precision=nb.PREC_FP32,
                                          max_depth=4, rr_start_depth=3, environment_albedo=albedo)

    def q(r):
        return orc.hybrid_query(setup["net"], setup["nbvh"], setup["ev"], setup["et"], setup["enodes"],
                                setup["eprims"], r, mode=orc.FP32)
    ref, rc = orc.path_trace(q, nb.camera_array(cam), W, H, max_depth=4, rr_start_depth=3, env_albedo=albedo,
                             verts=setup["ev"], tris=setup["et"])
    compare(img, ref, counts[0], rc)


def test_triangle_rendering_of_the_object_matches_oracle(nb, orc, setup):
    """The reference's ignore_neural_blas rendering (the object's own triangles, one albedo): the
    image the validation mode compares the neural rendering with."""
    cam = nb.camera_look_at((0.4, 0.3, 2.8), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    albedo = (0.9, 0.9, 0.9)
    img, counts = setup["mod"].path_trace(cam, W, H, triangles=True, max_depth=4, rr_start_depth=3,
                                          environment_albedo=albedo)

    def q(r):
        return orc.tri_query(setup["v"], setup["t"], setup["nodes"], setup["prims"], r)
    ref, rc = orc.path_trace(q, nb.camera_array(cam), W, H, max_depth=4, rr_start_depth=3, env_albedo=albedo,
                             verts=setup["v"], tris=setup["t"])
    compare(img, ref, counts[0], rc)
    assert counts[0][1] > 0  # bounces off the triangles
    with pytest.raises(nb.NbvhError):  # an environment is not combined with the triangle object
        setup["mod"].path_trace(cam, W, H, triangles=True, environment=setup["env"])


def test_secondary_visibility_threshold(nb, orc, setup):
    """primary_visibility_threshold at depth 1, secondary_visibility_threshold after
    (neural_bvh_module_common.cuh:171-173)."""
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    img, counts = setup["mod"].path_trace(cam, W, H, precision=nb.PREC_FP32, max_depth=3,
                                          visibility_threshold=0.5, secondary_visibility_threshold=0.500004)

    def q(r, depth):
        thr = 0.5 if depth == 1 else 0.500004
        return setup["net"].neural_query(setup["nbvh"], r, mode=orc.FP32, threshold=thr)
    ref, rc = orc.path_trace(q, nb.camera_array(cam), W, H, max_depth=3, query_takes_depth=True)
    compare(img, ref, counts[0], rc)
    base, bc = setup["mod"].path_trace(cam, W, H, precision=nb.PREC_FP32, max_depth=3)
    # the bounce threshold changes what the bounces hit, not the primary rays
    assert bc[0][:2] == counts[0][:2] and bc[0][2] != counts[0][2]


def test_path_trace_deterministic_and_accumulating(nb, setup):
    from nbvh_amd import hip
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    mod = setup["mod"]
    acc = hip.Buffer(W * H * 16)
    for prec in (nb.PREC_FP16, nb.PREC_FP32):
        mod.path_trace_device(cam, W, H, acc.ptr, precision=prec, sample=0, rr_start_depth=2)
        hip.synchronize()
        a0 = acc.download(np.float32)
        mod.path_trace_device(cam, W, H, acc.ptr, precision=prec, sample=0, rr_start_depth=2)
        hip.synchronize()
        assert np.array_equal(acc.download(np.float32), a0)  # same frame, bit-identical
        mod.path_trace_device(cam, W, H, acc.ptr, precision=prec, sample=1, rr_start_depth=2)
        hip.synchronize()
        a01 = acc.download(np.float32).reshape(-1, 4)
        acc.zero()
        mod.path_trace_device(cam, W, H, acc.ptr, precision=prec, sample=1, rr_start_depth=2)
        hip.synchronize()
        a1 = acc.download(np.float32).reshape(-1, 4)
        assert np.array_equal(a01[:, 3], np.full(W * H, 2, np.float32))
        assert np.array_equal(a01[:, :3], a0.reshape(-1, 4)[:, :3] + a1[:, :3])


@pytest.mark.parametrize("hybrid", [False, True])
def test_path_trace_graph_replays_match_eager(nb, setup, hybrid):
    """A frame recorded into one HIP graph (live-path counts on the device, buffers in a caller
    workspace, sorted bounces in bucket order) replays bit-identical to the eager frame, with eager
    frames and synchronisations between the replays; the eager frame with a workspace equals the
    one with transient memory and the one whose bounces read their counts back (wavefront rounds)."""
    from nbvh_amd import hip
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    mod, env = setup["mod"], (setup["env"] if hybrid else None)
    kw = dict(precision=nb.PREC_FP16, max_depth=4, rr_start_depth=3, sample=0)
    p0 = mod.path_params(W, H, **kw)
    size = mod.path_trace_workspace_size(p0, W, H, env)
    assert size >= W * H * 150
    ws = hip.Buffer(size)
    p = mod.path_params(W, H, workspace=ws, **kw)
    s = hip.Stream()
    acc_e, acc_g, acc_x = hip.Buffer(W * H * 16), hip.Buffer(W * H * 16), hip.Buffer(W * H * 16)
    counts = mod.path_trace_params_device(p, cam, W, H, acc_e.ptr, env, stream=s.handle)
    s.synchronize()
    eager = acc_e.download(np.float32)
    assert counts[0] == W * H and counts[1] > 0 and np.all(eager.reshape(-1, 4)[:, 3] == 1)
    assert mod.path_trace_params_device(p0, cam, W, H, acc_x.ptr, env, stream=s.handle) == counts
    s.synchronize()
    assert np.array_equal(acc_x.download(np.float32), eager)
    pw = mod.path_params(W, H, wavefront_bounces=True, **kw)
    assert mod.path_trace_params_device(pw, cam, W, H, acc_x.ptr, env, stream=s.handle) == counts
    s.synchronize()
    assert np.array_equal(acc_x.download(np.float32), eager)
    with hip.Graph.capture(s) as g:
        mod.path_trace_params_device(p, cam, W, H, acc_g.ptr, env, stream=s.handle, stats=False)
    try:
        for k in range(4):
            acc_g.zero()
            g.launch(s)
            s.synchronize()
            assert np.array_equal(acc_g.download(np.float32), eager), k
            # eager work between the replays: a frame of another sample into another buffer
            mod.path_trace_params_device(mod.path_params(W, H, **dict(kw, sample=0, rr_start_depth=2)), cam, W, H,
                                         acc_x.ptr, env, stream=s.handle)
            s.synchronize()
    finally:
        g.destroy()
        s.destroy()


def test_path_trace_workspace_errors(nb, setup):
    from nbvh_amd import hip
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    mod = setup["mod"]
    acc = hip.Buffer(W * H * 16)
    p0 = mod.path_params(W, H, precision=nb.PREC_FP16)
    size = mod.path_trace_workspace_size(p0, W, H)
    small = hip.Buffer(size - 256)
    with pytest.raises(nb.NbvhError):
        mod.path_trace_params_device(mod.path_params(W, H, precision=nb.PREC_FP16, workspace=small), cam, W, H,
                                     acc.ptr)
    s = hip.Stream()
    ws = hip.Buffer(size)
    try:
        with pytest.raises(nb.NbvhError):  # no workspace under capture
            with hip.Graph.capture(s):
                mod.path_trace_params_device(p0, cam, W, H, acc.ptr, stream=s.handle, stats=False)
        with pytest.raises(nb.NbvhError):  # stats read back under capture
            with hip.Graph.capture(s):
                mod.path_trace_params_device(mod.path_params(W, H, precision=nb.PREC_FP16, workspace=ws), cam, W, H,
                                             acc.ptr, stream=s.handle)
        # the stream is usable afterwards
        mod.path_trace_params_device(mod.path_params(W, H, precision=nb.PREC_FP16, workspace=ws), cam, W, H, acc.ptr,
                                     stream=s.handle)
        s.synchronize()
    finally:
        s.destroy()


def test_path_trace_fp16_close_to_fp32(nb, setup):
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    a32, _ = setup["mod"].path_trace(cam, W, H, precision=nb.PREC_FP32, samples=4)
    a16, _ = setup["mod"].path_trace(cam, W, H, precision=nb.PREC_FP16, samples=4)
    assert np.array_equal(a16[..., 3], np.full((H, W), 4, np.float32))
    assert abs(float(a16[..., :3].mean() - a32[..., :3].mean())) < 0.01 * float(a32[..., :3].mean())


def test_path_trace_errors(nb, setup):
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    with pytest.raises(nb.NbvhError):
        setup["mod"].path_trace(cam, W, H, max_depth=0)
    with pytest.raises(nb.NbvhError):
        setup["mod"].path_trace(cam, W, H, max_depth=17)
    img, counts = setup["mod"].path_trace(cam, 0, 0)
    assert img.size == 0 and counts == [[0] * 5]


def test_cpp_driver_path_mode_matches_python(nb, setup, tmp_path):
    """The C++ host driver's path mode (include/nbvh.hpp -> nbvh_path_trace) renders the same
    frames as the Python host path: 4 spp, max depth 5, roulette from depth 3, FP16."""
    import os
    import subprocess
    from conftest import ROOT
    out = tmp_path / "frame.ppm"
    r = subprocess.run([os.path.join(ROOT, "nbvh_amd", "nbvh_driver"), "path", "--width", "128", "--height", "96",
                        "--spp", "4", "--out", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    header = b"P6\n128 96\n255\n"
    assert data.startswith(header)
    got = np.frombuffer(data[len(header):], np.uint8).reshape(96, 128, 3).astype(np.int32)
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, 128, 96)
    img, _ = setup["mod"].path_trace(cam, 128, 96, samples=4, precision=nb.PREC_FP16, max_depth=5, rr_start_depth=3)
    v = np.clip(img[..., :3] / img[..., 3:4], 0, 1) ** (1 / 2.2)
    exp = np.floor(v * 255 + 0.5).astype(np.int32)
    assert np.abs(got - exp).max() <= 1
    assert got.mean() > 20


def test_path_trace_lod_switch_on_bounces(nb, orc, ctx):
    """path_lod_switch: rays at depth > 1 traverse the N-BVH at LoD lod_switch. An N-BVH of the
    benchmark mesh cut at 2048 leaves whose interior nodes at depth 5 are also marked LEAF_LOD(6)
    (children kept, as the reference's LoD assignment leaves them): primary rays at LoD 0, bounces
    at LoD 6, FP32, against the oracle's path trace with the same per-depth LoD; the frame differs
    from bounces at LoD 0."""
    v, t, nodes, prims, _ = bench_scene(nb)
    nbvh = nb.nbvh_cut(nodes, 2048)
    depth = np.full(len(nbvh), -1)
    depth[0] = 0
    marked = 0
    for i in range(len(nbvh)):
        if nbvh["prim_count"][i] == 0 and nbvh["left_or_prim_offset"][i] != nb.NO_NEURAL_CHILD:
            c = int(nbvh["left_or_prim_offset"][i])
            depth[c] = depth[c + 1] = depth[i] + 1
            if depth[i] == 5:
                nbvh["prim_count"][i] = 8  # LEAF_LOD(6), children kept
                marked += 1
    assert marked > 8 and np.all(depth >= 0)
    scene = nb.Scene(ctx, v, t)
    scene.set_nbvh(nbvh)
    scene.init_network(nb.network_config(), seed=4, table_range=1e-3)
    cfg, tables, mlp = scene.network()
    lo, hi = scene.bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    nodes_d = scene.nbvh()
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    mod = nb.NeuralBVHModule(ctx, scene)
    img, counts = mod.path_trace(cam, W, H, precision=nb.PREC_FP32, max_depth=4, bounce_lod=6)
    ref, rc = orc.path_trace(lambda r, d: net.neural_query(nodes_d, r, lod=6 if d > 1 else 0, mode=orc.FP32),
                             nb.camera_array(cam), W, H, max_depth=4, query_takes_depth=True)
    compare(img, ref, counts[0], rc)
    assert counts[0][1] > 100
    img0, _ = mod.path_trace(cam, W, H, precision=nb.PREC_FP32, max_depth=4)
    assert not np.array_equal(img0, img)
    scene.close()
