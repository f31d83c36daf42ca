"""nbvh_query_noalloc, the allocation-free query, against nbvh_query through the C ABI.

Its scratch (queue counters, sort keys and claim order) is one static array per GPU that every
call reuses, so each case runs two calls in a row, the second on a smaller input, and each must
give nbvh_query's hit records bit for bit. Also: the documented limits are refused before any
launch.
"""
import ctypes

import numpy as np
import pytest

from conftest import bench_scene, camera_rays, random_rays

pytestmark = pytest.mark.gpu


def hits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def setup(nb):
    ctx = nb.Context()
    v, t, _, _, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(nb.network_config(), seed=1, table_range=1e-4)
    yield ctx, scene, nb.NeuralBVHModule(ctx, scene)
    scene.close()
    ctx.close()


def run_pair(nb, mod, params, rays):
    """(nbvh_query hits, nbvh_query_noalloc hits) of the same device rays."""
    from nbvh_amd import hip
    dr = hip.Buffer.from_array(rays)
    ha, hb = hip.Buffer(len(rays) * 48), hip.Buffer(len(rays) * 48)
    hb.zero()
    s = hip.Stream()
    mod.query_device(params, dr.ptr, len(rays), ha.ptr, stream=s.handle)
    s.synchronize()
    mod.query_device_noalloc(params, dr.ptr, len(rays), hb.ptr, stream=s.handle)
    s.synchronize()
    s.destroy()
    return ha.download(nb.HIT_DTYPE), hb.download(nb.HIT_DTYPE)


@pytest.mark.parametrize("precision,kind,sort", [
    ("PREC_FP16", "QUERY_NEURAL_CLOSEST", True),
    ("PREC_FP16", "QUERY_NEURAL_ANY", False),
    ("PREC_FP32", "QUERY_NEURAL_CLOSEST", False),
    ("PREC_FP16", "QUERY_TRIANGLE_CLOSEST", True),
])
def test_noalloc_matches_nbvh_query_on_two_calls(nb, orc, setup, precision, kind, sort):
    _, _, mod = setup
    p = mod.params(kind=getattr(nb, kind), precision=getattr(nb, precision), sort_rays=sort)
    big = random_rays(500_000 if sort else 700_000, 41, inside_frac=0.3)
    _, cam = camera_rays(nb, orc, 160, 120)
    small = np.concatenate([cam, random_rays(3_001, 42, inside_frac=0.5)])
    for i, rays in enumerate((big, small)):
        print(f"noalloc case {precision} {kind} sort={sort} call {i} n={len(rays)}", flush=True)
        ref, got = run_pair(nb, mod, p, rays)
        assert hits_equal(got, ref), (i, len(rays))
        assert (nb.hit_kind(got) != 0).sum() > 0


def test_noalloc_refuses_its_limits(nb, setup):
    from nbvh_amd import _ffi, hip
    ctx, scene, mod = setup
    L = _ffi.lib()
    n = _ffi.NOALLOC_MAX_SORTED_RAYS + 1
    dr, dh = hip.Buffer(n * 32), hip.Buffer(n * 48)
    p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16, sort_rays=True)
    st = L.nbvh_query_noalloc(ctx.handle, scene.handle, ctypes.byref(p), ctypes.c_void_p(dr.ptr), n,
                              ctypes.c_void_p(dh.ptr), None)
    assert st == _ffi.ERR_CAPACITY
    assert b"at most 524288 rays" in L.nbvh_last_error(ctx.handle)
    p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=nb.PREC_FP16, wavefront=True)
    st = L.nbvh_query_noalloc(ctx.handle, scene.handle, ctypes.byref(p), ctypes.c_void_p(dr.ptr), 1000,
                              ctypes.c_void_p(dh.ptr), None)
    assert st == _ffi.ERR_UNSUPPORTED
