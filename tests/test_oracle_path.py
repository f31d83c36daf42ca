"""Pins of the path-trace restatement (oracle.path_trace / oracle_path_shade) by closed forms.

A scripted query stands in for the scene: it answers every ray with a neural hit at t = 1 with a
fixed normal and albedo up to a given depth and with a miss afterwards. Then the radiance of a
pixel has a closed form: the product of the albedos along the path times the background of the
last ray, and the bounce directions have the moments of a cosine-weighted hemisphere.
"""
import numpy as np
import pytest


def scripted_query(orc, hit_depths, albedo, normal=(0.0, 0.0, 1.0), log=None):
    """query(rays) for oracle.path_trace: neural hits at t = 1 for the first `hit_depths` calls,
    misses afterwards. Records each call's rays in `log`."""
    calls = [0]

    def q(rays):
        calls[0] += 1
        if log is not None:
            log.append(rays.copy())
        hits = np.zeros(len(rays), orc.HIT_DTYPE)
        if calls[0] <= hit_depths:
            hits["info"] = 1
            hits["t"] = 1.0
            hits["normal_or_uv"] = normal
            hits["albedo"] = albedo
        else:
            hits["t"] = np.float32(3.4028234663852886e38)
        return hits
    return q


def sky(dy):
    a = 0.5 * (dy.astype(np.float64) + 1.0)
    return np.stack([(1 - a) + a * 0.5, (1 - a) + a * 0.7, (1 - a) + a * 1.0], axis=1)


@pytest.fixture(scope="module")
def cam(nb):
    return nb.camera_array(nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, 64, 48))


def test_escaped_primary_rays_see_the_background(orc, cam):
    log = []
    acc, counts = orc.path_trace(scripted_query(orc, 0, (0.5,) * 3, log=log), cam, 64, 48, background=1)
    assert counts == [64 * 48]
    assert np.array_equal(acc[:, 3], np.ones(64 * 48, np.float32))
    np.testing.assert_allclose(acc[:, :3], sky(log[0]["d"][:, 1]), rtol=0, atol=2e-7)
    acc, _ = orc.path_trace(scripted_query(orc, 0, (0.5,) * 3), cam, 64, 48, background=0,
                            background_color=(0.25, 0.5, 2.0))
    assert np.array_equal(acc[:, :3], np.tile(np.float32([0.25, 0.5, 2.0]), (64 * 48, 1)))


@pytest.mark.parametrize("hit_depths", [1, 2, 3])
def test_throughput_is_the_product_of_albedos(orc, cam, hit_depths):
    albedo = np.float32([0.5, 0.25, 0.75])
    log = []
    acc, counts = orc.path_trace(scripted_query(orc, hit_depths, albedo, log=log), cam, 64, 48, max_depth=5,
                                 background=1)
    assert counts == [64 * 48] * (hit_depths + 1)
    last = log[hit_depths]
    # Paths are appended in order, so the live arrays keep pixel order here.
    expect = sky(last["d"][:, 1]) * albedo.astype(np.float64) ** hit_depths
    np.testing.assert_allclose(acc[:, :3], expect, rtol=2e-6, atol=0)
    # The next origin is the hit point offset by 1e-4 along the normal turned toward the ray.
    prev = log[hit_depths - 1]
    n = np.where(prev["d"][:, 2:3] > 0, -1.0, 1.0) * np.array([0.0, 0.0, 1.0])
    np.testing.assert_allclose(last["o"], prev["o"] + prev["d"] + 1e-4 * n, rtol=0, atol=2e-6)


def test_paths_end_at_max_depth(orc, cam):
    acc, counts = orc.path_trace(scripted_query(orc, 10, (0.9,) * 3), cam, 64, 48, max_depth=3)
    assert counts == [64 * 48] * 3
    assert np.array_equal(acc[:, :3], np.zeros((64 * 48, 3), np.float32))
    assert np.array_equal(acc[:, 3], np.ones(64 * 48, np.float32))


def test_bounce_directions_are_cosine_distributed(orc, nb):
    w, h = 256, 256
    cam = nb.camera_array(nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, w, h))
    log = []
    orc.path_trace(scripted_query(orc, 1, (1.0,) * 3, log=log), cam, w, h, max_depth=3)
    d = log[1]["d"].astype(np.float64)
    # Primary rays travel toward -z, so the shading normal stays +z; cos theta = d.z.
    c = d[:, 2]
    assert np.all(c >= 0) and np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    n = len(c)
    assert abs(c.mean() - 2 / 3) < 4 * np.sqrt(1 / 18 / n)       # E[cos] = 2/3, Var = 1/18
    assert abs((c ** 2).mean() - 1 / 2) < 4 * np.sqrt(1 / 12 / n)  # E[cos^2] = 1/2, Var = 1/12
    phi = np.arctan2(d[:, 1], d[:, 0])
    assert abs(np.cos(phi).mean()) < 0.02 and abs(np.sin(phi).mean()) < 0.02


def test_russian_roulette_is_unbiased(orc, nb):
    w, h = 256, 256
    cam = nb.camera_array(nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, w, h))
    log = []
    acc, counts = orc.path_trace(scripted_query(orc, 2, (0.5,) * 3, log=log), cam, w, h, max_depth=5,
                                 rr_start_depth=2, background=0, background_color=(1.0, 1.0, 1.0))
    # Depth 2: throughput 0.25 survives with p = 0.25 and becomes 1.
    n = w * h
    alive = counts[2] / n
    assert abs(alive - 0.25) < 4 * np.sqrt(0.25 * 0.75 / n)
    L = acc[:, 0].astype(np.float64)
    assert set(np.unique(L)) <= {0.0, 1.0}
    assert abs(L.mean() - 0.25) < 4 * np.sqrt(0.25 * 0.75 / n)
    assert np.array_equal(acc[:, 3], np.ones(n, np.float32))


def test_samples_accumulate(orc, cam):
    a0, _ = orc.path_trace(scripted_query(orc, 1, (0.5,) * 3), cam, 64, 48, sample=0)
    acc = a0.copy()
    orc.path_trace(scripted_query(orc, 1, (0.5,) * 3), cam, 64, 48, sample=1, accum=acc)
    a1 = np.zeros_like(a0)
    orc.path_trace(scripted_query(orc, 1, (0.5,) * 3), cam, 64, 48, sample=1, accum=a1)
    assert np.array_equal(acc[:, 3], np.full(64 * 48, 2, np.float32))
    assert np.array_equal(acc[:, :3], a0[:, :3] + a1[:, :3])
    assert not np.array_equal(a0[:, :3], a1[:, :3])
