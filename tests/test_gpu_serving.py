"""GPU tests of the continuous-batching request loop (nbvh_server_* through the C ABI).

Bar: every request's hit records are bit-identical to the same rays traced by a direct
nbvh_query, alone or inside a mixed stream of concurrent requests of other sizes, kinds and
scenes (each ray's result depends on that ray alone).
"""
import threading

import numpy as np
import pytest

from conftest import bench_scene, camera_rays, random_rays

pytestmark = pytest.mark.gpu


def hits_equal(a, b):
    return np.array_equal(a.view(np.uint32).reshape(len(a), -1), b.view(np.uint32).reshape(len(b), -1))


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def scenes(nb, ctx):
    v, t, *_ = bench_scene(nb)
    a = nb.Scene(ctx, v, t)
    a.build_nbvh(256)
    a.init_network(nb.network_config(), seed=1, table_range=1e-4)
    v2, t2 = nb.icosphere(4, seed=3, amplitude=0.2, center=(0.3, -0.2, 0.1), radius=0.7)
    b = nb.Scene(ctx, v2, t2)
    b.build_nbvh(64)
    b.init_network(nb.network_config(), seed=5, table_range=1e-2)
    yield a, b
    a.close()
    b.close()


def direct(nb, ctx, scene, p, rays):
    m = nb.NeuralBVHModule(ctx, scene)
    return m.neural_bvh_traversal(rays, kind=p.kind, precision=p.precision, lod=p.lod,
                                  max_traversal_steps=p.max_traversal_steps,
                                  visibility_threshold=p.visibility_threshold)


def test_host_requests_match_direct_query(nb, orc, ctx, scenes):
    a, _ = scenes
    _, rays = camera_rays(nb, orc, 64, 48)
    rays = np.concatenate([rays, random_rays(2000, 21, inside_frac=0.3)])
    with nb.Server(ctx) as srv:
        for kind, prec in [(nb.QUERY_NEURAL_CLOSEST, nb.PREC_FP16), (nb.QUERY_NEURAL_CLOSEST, nb.PREC_FP32),
                           (nb.QUERY_NEURAL_ANY, nb.PREC_BF16), (nb.QUERY_TRIANGLE_ANY, nb.PREC_FP32)]:
            p = nb.NeuralBVHModule.params(kind=kind, precision=prec)
            hits, info = srv.wait(srv.submit(a, p, rays))
            assert info["status"] == 0 and info["n_launches"] == 1
            assert info["submit_ns"] <= info["launch_ns"] <= info["complete_ns"]
            assert hits_equal(hits, direct(nb, ctx, a, p, rays)), (kind, prec)


def test_alone_and_in_the_mix_are_identical(nb, orc, ctx, scenes):
    a, b = scenes
    _, cam = camera_rays(nb, orc, 96, 64)
    probe = np.concatenate([cam, random_rays(3000, 31, inside_frac=0.4)])
    p16 = nb.NeuralBVHModule.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16)
    with nb.Server(ctx, n_streams=3, max_batch_rays=60000) as srv:
        alone, info = srv.wait(srv.submit(a, p16, probe))
        assert info["batch_parts"] == 1
    assert hits_equal(alone, direct(nb, ctx, a, p16, probe))

    # The mix: three client threads submit requests of uneven sizes against both scenes with
    # four kinds/precisions; the probe request goes in the middle of one thread's stream.
    rng = np.random.default_rng(7)
    kinds = [(nb.QUERY_NEURAL_CLOSEST, nb.PREC_FP16), (nb.QUERY_NEURAL_ANY, nb.PREC_FP16),
             (nb.QUERY_NEURAL_CLOSEST, nb.PREC_FP32), (nb.QUERY_TRIANGLE_CLOSEST, nb.PREC_FP32)]
    jobs = []
    for j in range(3):
        mine = []
        for k in range(14):
            n = int(np.clip(rng.lognormal(np.log(3000), 1.0), 64, 30000))
            kind, prec = kinds[rng.integers(len(kinds))]
            sc = (a, b)[rng.integers(2)]
            mine.append((sc, nb.NeuralBVHModule.params(kind=kind, precision=prec),
                         random_rays(n, 1000 * j + k, inside_frac=0.3)))
        if j == 1:
            mine.insert(7, (a, p16, probe))
        jobs.append(mine)
    results = [[None] * len(m) for m in jobs]
    with nb.Server(ctx, n_streams=3, max_batch_rays=60000) as srv:
        def client(j):
            tickets = [srv.submit(sc, p, r) for sc, p, r in jobs[j]]
            for i, t in enumerate(tickets):
                results[j][i] = srv.wait(t)
        threads = [threading.Thread(target=client, args=(j,)) for j in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        stats = srv.stats()
    assert stats["n_completed"] == sum(len(m) for m in jobs) and stats["n_failed"] == 0
    assert stats["n_batched_requests"] > stats["n_launches"]  # requests shared launches
    mixed, info = results[1][7]
    assert hits_equal(mixed, alone)
    for j, mine in enumerate(jobs):
        for i, (sc, p, r) in enumerate(mine):
            assert hits_equal(results[j][i][0], direct(nb, ctx, sc, p, r)), (j, i)


def test_device_requests(nb, orc, ctx, scenes):
    from nbvh_amd import hip
    a, _ = scenes
    rays = random_rays(5000, 41, inside_frac=0.3)
    p = nb.NeuralBVHModule.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16)
    dr = hip.Buffer.from_array(rays)
    dh = [hip.Buffer(len(rays) * 48) for _ in range(3)]
    with nb.Server(ctx) as srv:
        tickets = [srv.submit_device(a, p, dr.ptr, len(rays), d.ptr) for d in dh]
        for t in tickets:
            hits, info = srv.wait(t)
            assert hits is None and info["status"] == 0
    exp = direct(nb, ctx, a, p, rays)
    for d in dh:
        assert hits_equal(d.download(nb.HIT_DTYPE), exp)


def test_staging_admission_and_capacity(nb, ctx, scenes):
    a, _ = scenes
    p = nb.NeuralBVHModule.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16)
    # 1 MiB arena: two 6000-ray requests (480 KB each) fit at once, the rest wait for space.
    with nb.Server(ctx, staging_bytes=1 << 20) as srv:
        batches = [random_rays(6000, 50 + i, inside_frac=0.3) for i in range(6)]
        tickets = [srv.submit(a, p, r) for r in batches]
        for t, r in zip(tickets, batches):
            hits, info = srv.wait(t)
            assert info["batch_parts"] <= 2
            assert hits_equal(hits, direct(nb, ctx, a, p, r))
        with pytest.raises(nb.NbvhError) as e:
            srv.submit(a, p, random_rays(20000, 60))
        assert e.value.status == 9  # larger than the arena


def test_timeout_unknown_ticket_and_failed_request(nb, ctx, scenes):
    a, _ = scenes
    p = nb.NeuralBVHModule.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP32)
    with nb.Server(ctx) as srv:
        t = srv.submit(a, p, random_rays(200000, 70, inside_frac=0.3))
        try:
            srv.wait(t, timeout_us=0)
            done = True
        except nb.NbvhError as e:
            assert e.status == 8
            done = False
        if not done:
            hits, info = srv.wait(t)  # the ticket survived the timeout
            assert info["status"] == 0
        with pytest.raises(nb.NbvhError) as e:
            srv.wait(t)
        assert e.value.status == 1  # collected already
        empty = nb.Scene(ctx)
        t = srv.submit(empty, p, random_rays(100, 71))
        with pytest.raises(nb.NbvhError) as e:
            srv.wait(t)
        assert e.value.status == 6  # no N-BVH: the request fails, the server goes on
        hits, info = srv.wait(srv.submit(a, p, random_rays(100, 72)))
        assert info["status"] == 0
        empty.close()


def test_uniform_segments_alone_and_in_the_mix(nb, orc, ctx):
    """A network with the uniform segment distribution: each ray's stream is seeded by its index in
    its own request, so a request's hits are the same alone, batched with other requests of the
    same seed, and traced by a direct query; requests of other seeds are not batched with it."""
    v, t, *_ = bench_scene(nb)
    s = nb.Scene(ctx, v, t)
    s.build_nbvh(256)
    s.init_network(nb.network_config(segment_distribution=nb.SEGMENT_UNIFORM), seed=7, table_range=1e-1)
    m = nb.NeuralBVHModule(ctx, s)
    reqs = [random_rays(n, 40 + i, inside_frac=0.3) for i, n in enumerate((3000, 1200, 2500, 800))]
    seeds = [11, 11, 11, 12]
    exp = [m.neural_bvh_traversal(r, precision=nb.PREC_FP32, segment_seed=sd) for r, sd in zip(reqs, seeds)]
    with nb.Server(ctx) as srv:
        alone = [srv.wait(srv.submit(s, nb.NeuralBVHModule.params(precision=nb.PREC_FP32, segment_seed=sd), r))
                 for r, sd in zip(reqs, seeds)]
        tickets = [srv.submit(s, nb.NeuralBVHModule.params(precision=nb.PREC_FP32, segment_seed=sd), r)
                   for r, sd in zip(reqs, seeds)]
        mixed = [srv.wait(tk) for tk in tickets]
    for (ha, ia), (hm, im), e in zip(alone, mixed, exp):
        assert ia["status"] == 0 and im["status"] == 0
        assert hits_equal(ha, e) and hits_equal(hm, e)
    s.close()
