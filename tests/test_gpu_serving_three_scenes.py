"""Continuous batching over three resident scenes at their full sizes (the serving workload):
the 81,920-triangle icosphere (N-BVH 256), a field of 64 displaced geodesic spheres (2.17M
triangles, N-BVH 4096) and one geodesic sphere of frequency 648 (8.4M triangles, N-BVH 16384),
each with its own seeded FP16 network of the headline shape.

A seeded stream of requests, lognormal sizes (median 65,536 rays, clamped to [4,096, 2M]),
70 % closest hit / 30 % shadow any hit, scene drawn uniformly, half over host buffers (staged
through the server's arena) and half over device buffers, is submitted by three client threads
with exponential gaps between submissions. Bars:
  - every request's hit records are bit-identical to a direct query of the same rays;
  - a probe request is bit-identical alone and inside the mix;
  - on a sample of each scene's rays, the FP32 query is bit-exact with the oracle's FP32 path and
    the FP16 query's hit/miss differs from the FP64 oracle on at most 0.5 % of the rays.
"""
import threading
import time

import numpy as np
import pytest

from conftest import bench_scene

pytestmark = pytest.mark.gpu

W, H = 2048, 1024


def hits_equal(a, b):
    return np.array_equal(a.view(np.uint32).reshape(len(a), -1), b.view(np.uint32).reshape(len(b), -1))


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def three(nb, ctx):
    """The three scenes and, per scene, a 2048x1024 camera frame of rays (device and host copies)."""
    from nbvh_amd import hip
    v, t, *_ = bench_scene(nb)
    specs = [((v, t), 256, 1), (nb.sphere_field(64, seed=5), 4096, 2), (nb.geosphere(648, seed=9, amplitude=0.1), 16384, 3)]
    out = []
    for (verts, tris), leaves, seed in specs:
        sc = nb.Scene(ctx, verts, tris)
        sc.build_nbvh(leaves)
        sc.init_network(nb.network_config(), seed=seed, table_range=1e-4)
        lo, hi = sc.bbox()
        c = (lo + hi) * 0.5
        ext = float(np.max(hi - lo))
        cam = nb.camera_look_at(tuple(c + np.array([0.3, 0.4, 1.0]) * 1.6 * ext), tuple(c), (0, 1, 0), 45.0, W, H)
        d = hip.Buffer(W * H * 32)
        ctx.raygen(cam, W, H, d.ptr)
        hip.synchronize()
        out.append(dict(scene=sc, tris=len(tris), d_rays=d, rays=d.download(nb.RAY_DTYPE),
                        mod=nb.NeuralBVHModule(ctx, sc)))
    assert [o["tris"] for o in out][0] == 81920 and out[1]["tris"] > 2_000_000 and out[2]["tris"] > 8_000_000
    assert [len(o["scene"].nbvh()) for o in out] == [511, 8191, 32767]
    yield out
    for o in out:
        o["scene"].close()


def _requests(nb, n_req, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_req):
        n = int(np.clip(rng.lognormal(np.log(65536), 1.0), 4096, 2 << 20))
        sc = int(rng.integers(3))
        kind = nb.QUERY_NEURAL_CLOSEST if rng.random() < 0.7 else nb.QUERY_NEURAL_ANY
        off = int(rng.integers(0, W * H - n + 1))
        reqs.append(dict(scene=sc, kind=kind, off=off, n=n, device=bool(i % 2), gap=float(rng.exponential(5e-4))))
    return reqs


def test_three_scene_mix(nb, orc, ctx, three):
    from nbvh_amd import hip
    p16 = {k: nb.NeuralBVHModule.params(kind=k, precision=nb.PREC_FP16)
           for k in (nb.QUERY_NEURAL_CLOSEST, nb.QUERY_NEURAL_ANY)}
    jobs = [_requests(nb, 16, 100 + j) for j in range(3)]
    probe = dict(scene=2, kind=nb.QUERY_NEURAL_CLOSEST, off=W * H // 2 - 32768, n=65536, device=False, gap=0.0)
    jobs[1].insert(8, probe)
    n_rays = sum(r["n"] for m in jobs for r in m)
    print(f"[three-scene mix] {sum(len(m) for m in jobs)} requests, {n_rays} rays", flush=True)

    def rays_of(r):
        return three[r["scene"]]["rays"][r["off"]:r["off"] + r["n"]]

    with nb.Server(ctx) as srv:
        alone, info = srv.wait(srv.submit(three[2]["scene"], p16[probe["kind"]], rays_of(probe)))
        assert info["status"] == 0 and info["batch_parts"] == 1
    dev_out = {}
    for j, mine in enumerate(jobs):
        for i, r in enumerate(mine):
            if r["device"]:
                dev_out[(j, i)] = hip.Buffer(r["n"] * 48)
    results = [[None] * len(m) for m in jobs]
    with nb.Server(ctx) as srv:
        def client(j):
            tickets = []
            for i, r in enumerate(jobs[j]):
                time.sleep(r["gap"])
                sc = three[r["scene"]]
                if r["device"]:
                    tickets.append(srv.submit_device(sc["scene"], p16[r["kind"]], sc["d_rays"].ptr + 32 * r["off"],
                                                     r["n"], dev_out[(j, i)].ptr))
                else:
                    tickets.append(srv.submit(sc["scene"], p16[r["kind"]], rays_of(r)))
            for i, t in enumerate(tickets):
                results[j][i] = srv.wait(t)

        threads = [threading.Thread(target=client, args=(j,)) for j in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        stats = srv.stats()
    assert stats["n_failed"] == 0 and stats["n_completed"] == sum(len(m) for m in jobs)
    assert stats["n_batched_requests"] > stats["n_launches"], stats  # requests shared launches
    assert hits_equal(results[1][8][0], alone)
    for j, mine in enumerate(jobs):
        for i, r in enumerate(mine):
            sc = three[r["scene"]]
            hits, info = results[j][i]
            assert info["status"] == 0
            if r["device"]:
                assert hits is None
                hits = dev_out[(j, i)].download(nb.HIT_DTYPE)
            d = hip.Buffer(r["n"] * 48)
            sc["mod"].query_device(p16[r["kind"]], sc["d_rays"].ptr + 32 * r["off"], r["n"], d.ptr)
            hip.synchronize()
            assert hits_equal(hits, d.download(nb.HIT_DTYPE)), (j, i, r)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_scene_samples_against_oracle(nb, orc, three, which):
    """FP32 bit-exact and FP16 hit/miss against FP64 on every 1,399th ray of the scene's frame."""
    sc = three[which]
    cfg, tables, mlp = sc["scene"].network()
    lo, hi = sc["scene"].bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    sample = sc["rays"][::1399]
    g32 = sc["mod"].neural_bvh_traversal(sample, precision=nb.PREC_FP32)
    assert hits_equal(g32, net.neural_query(sc["scene"].nbvh(), sample, mode=orc.FP32))
    e64 = net.neural_query(sc["scene"].nbvh(), sample, mode=orc.FP64)
    g16 = sc["mod"].neural_bvh_traversal(sample, precision=nb.PREC_FP16)
    kg, ke = nb.hit_kind(g16), nb.hit_kind(e64)
    assert (ke != 0).mean() > 0.05  # the frame sees the scene
    assert (kg != ke).mean() <= 0.005, (kg != ke).mean()
    a16 = sc["mod"].neural_bvh_traversal(sample, kind=nb.QUERY_NEURAL_ANY, precision=nb.PREC_FP16)
    a64 = net.neural_query(sc["scene"].nbvh(), sample, mode=orc.FP64, any_hit=True)
    assert (nb.hit_kind(a16) != nb.hit_kind(a64)).mean() <= 0.005
