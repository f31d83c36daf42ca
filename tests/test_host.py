"""Host side of libnbvh_hip.so (no GPU): ABI exports, builders, formats, error handling."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, GOLDEN


def header_functions():
    src = open(os.path.join(ROOT, "include", "nbvh.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|uint32_t|const char\*)\s+(nbvh_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol(nb):
    from nbvh_amd import _ffi
    names = header_functions()
    assert len(names) >= 30
    L = _ffi.lib()
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _ffi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (nbvh_\w+)", out))
    assert set(names) <= exported
    # Python binding covers the whole header.
    assert set(names) == set(_ffi.SIGNATURES)
    assert L.nbvh_abi_version() == 2


def test_status_strings_and_null_args(nb):
    from nbvh_amd import _ffi
    L = _ffi.lib()
    assert L.nbvh_status_string(0) == b"ok"
    assert L.nbvh_status_string(5) == b"format error"
    assert L.nbvh_bvh_build(None, 0, None, 0, 2, None, 0, None, None) == _ffi.ERR_INVALID_ARGUMENT
    assert b"NULL" in L.nbvh_last_error(None)


def test_icosphere_counts_and_displacement(nb):
    for s in range(4):
        v, t = nb.icosphere(s, seed=1, amplitude=0.0)
        assert len(t) == 20 * 4 ** s and len(v) == 10 * 4 ** s + 2
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    v, t = nb.icosphere(6, seed=7, amplitude=0.1)
    assert len(t) == 81920 and len(v) == 40962
    r = np.linalg.norm(v, axis=1)
    assert 0.9 <= r.min() < 0.97 and 1.03 < r.max() <= 1.1
    v2, _ = nb.icosphere(6, seed=7, amplitude=0.1)
    v3, _ = nb.icosphere(6, seed=8, amplitude=0.1)
    assert np.array_equal(v, v2) and not np.array_equal(v, v3)
    # Every edge is shared by exactly two triangles (closed, manifold).
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)


@pytest.mark.parametrize("k", [1, 2, 3, 7, 16])
def test_geosphere_closed_oriented_surface(nb, k):
    v, t = nb.geosphere(k, seed=1, amplitude=0.0)
    assert len(t) == 20 * k * k and len(v) == 10 * k * k + 2
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    assert len(np.unique(t)) == len(v)  # every vertex is used (shared edge points)
    # Each directed edge appears once and its reverse once: closed and consistently oriented.
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]).astype(np.int64)
    key = e[:, 0] * len(v) + e[:, 1]
    rev = e[:, 1] * len(v) + e[:, 0]
    assert len(np.unique(key)) == len(key) and np.array_equal(np.sort(key), np.sort(rev))
    p = v[t].astype(np.float64)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, p.mean(axis=1)) > 0)  # outward
    # k = 1 is the icosahedron itself.
    if k == 1:
        v0, t0 = nb.icosphere(0, amplitude=0.0)
        assert np.array_equal(v, v0) and np.array_equal(t, t0)


def test_sphere_field_workload(nb):
    v, t = nb.sphere_field(n_spheres=8, seed=3, freq_range=(4, 6))
    assert len(t) % 20 == 0 and 8 * 20 * 16 <= len(t) <= 8 * 20 * 36
    assert v.min() >= 0.0 and v.max() <= 1.0 + 1e-6
    assert t.max() == len(v) - 1
    v2, t2 = nb.sphere_field(n_spheres=8, seed=3, freq_range=(4, 6))
    assert np.array_equal(v, v2) and np.array_equal(t, t2)


def bvh_invariants(nodes, prims, v, t, max_leaf):
    n = len(nodes)
    assert sorted(prims.tolist()) == list(range(len(t)))
    seen = np.zeros(len(t), int)
    depth = np.zeros(n, int)
    stack = [0]
    visited = 0
    while stack:
        i = stack.pop()
        visited += 1
        nd = nodes[i]
        if nd["prim_count"] > 0:
            assert nd["prim_count"] <= max_leaf
            for p in prims[nd["left_or_prim_offset"]:nd["left_or_prim_offset"] + nd["prim_count"]]:
                seen[p] += 1
                tv = v[t[p]]
                assert np.all(tv >= nd["aabb_min"]) and np.all(tv <= nd["aabb_max"])
        else:
            l = nd["left_or_prim_offset"]
            assert i < l and l + 1 < n
            for c in (l, l + 1):
                assert np.all(nodes[c]["aabb_min"] >= nd["aabb_min"]) and np.all(nodes[c]["aabb_max"] <= nd["aabb_max"])
                depth[c] = depth[i] + 1
                stack.append(c)
    assert visited == n and np.all(seen == 1)
    return depth.max()


def test_bvh_build_invariants(nb):
    v, t = nb.icosphere(4, seed=2, amplitude=0.1)
    nodes, prims = nb.bvh_build(v, t)
    assert bvh_invariants(nodes, prims, v, t, 2) < 64
    # Degenerate input: all triangles identical (forces median splits).
    tt = np.tile(t[:1], (100, 1))
    nodes2, prims2 = nb.bvh_build(v, tt, max_leaf_size=2)
    bvh_invariants(nodes2, prims2, v, tt, 2)


def test_bvh_build_rejects_bad_input(nb):
    v, t = nb.icosphere(1)
    bad = t.copy()
    bad[3, 1] = 10 ** 6
    with pytest.raises(nb.NbvhError) as e:
        nb.bvh_build(v, bad)
    assert e.value.status == 1


def test_nbvh_cut_invariants(nb):
    v, t = nb.icosphere(5, seed=4, amplitude=0.1)
    nodes, prims = nb.bvh_build(v, t)
    for k in (1, 2, 64, 256):
        cut = nb.nbvh_cut(nodes, k)
        leaves = cut["left_or_prim_offset"] == nb.NO_NEURAL_CHILD
        assert leaves.sum() == k and len(cut) == 2 * k - 1
        assert set(cut["prim_count"][leaves].tolist()) <= {1, 2}
        assert np.all(cut["prim_count"][~leaves] == 0)
        assert np.array_equal(cut[0]["aabb_min"], nodes[0]["aabb_min"])
        # Every triangle (all three vertices) lies inside one leaf box.
        tmin, tmax = v[t].min(axis=1), v[t].max(axis=1)
        inside = np.zeros(len(t), bool)
        for lf in cut[leaves]:
            inside |= np.all((tmin >= lf["aabb_min"]) & (tmax <= lf["aabb_max"]), axis=1)
        assert inside.all()
        # Children adjacent, after the parent; each child's box inside its parent's.
        for i in np.nonzero(~leaves)[0]:
            l = cut[i]["left_or_prim_offset"]
            assert l > i and l + 1 < len(cut)
            for c in (l, l + 1):
                assert np.all(cut[c]["aabb_min"] >= cut[i]["aabb_min"])
                assert np.all(cut[c]["aabb_max"] <= cut[i]["aabb_max"])
    # Asking for more leaves than the BVH has gives the whole tree (mesh leaves: LEAF_FINAL).
    tiny_v, tiny_t = nb.icosphere(0)
    tn, _ = nb.bvh_build(tiny_v, tiny_t)
    full = nb.nbvh_cut(tn, 1000)
    assert len(full) == len(tn) and np.all(full["prim_count"][full["left_or_prim_offset"] == nb.NO_NEURAL_CHILD] == 1)


def test_nbvh_roundtrip_and_corruption(nb):
    data = open(os.path.join(GOLDEN, "chess_nodes_32940_hash_18_lvl_8.nbvh"), "rb").read()
    nodes = nb.nbvh_parse(data)
    assert nb.nbvh_serialize(nodes) == data
    for bad in (data[:-1], data[:15], data + b"\0", b""):
        with pytest.raises(nb.NbvhError) as e:
            nb.nbvh_parse(bad)
        assert e.value.status == 5
    # A child index pointing outside the array, or a cycle, is rejected.
    arr = bytearray(data)
    arr[16 + 24:16 + 32] = (10 ** 6).to_bytes(8, "little")
    with pytest.raises(nb.NbvhError):
        nb.nbvh_parse(bytes(arr))
    arr = bytearray(data)
    arr[16 + 40 + 24:16 + 40 + 32] = (0).to_bytes(8, "little")  # node 1 -> child 0 (cycle)
    with pytest.raises(nb.NbvhError):
        nb.nbvh_parse(bytes(arr))


def test_network_params(nb, orc):
    cfg = nb.network_config()
    nt, nm = nb.network_param_counts(cfg)
    assert nm == 18944 and nt == 2 * (4096 + 13824 + 46656 + 157464 + 12 * 2 ** 19)
    t, m = nb.network_init_params(cfg, 3, 1e-4)
    assert np.abs(t).max() <= 1e-4 and abs(t.mean()) < 2e-6
    assert np.abs(m[:64 * 96]).max() <= np.sqrt(6 / 96) and np.abs(m[64 * 96:64 * 96 + 4096]).max() <= np.sqrt(6 / 64)
    t2, m2 = nb.network_init_params(cfg, 3, 1e-4)
    assert np.array_equal(t, t2) and np.array_equal(m, m2)
    bad = nb.network_config(n_neurons=32)
    with pytest.raises(nb.NbvhError) as e:
        nb.network_param_counts(bad)
    assert e.value.status == 7
    # Chess configuration (8 levels x 4 features, base 8, scale 2, 2^18).
    chess = nb.network_config(n_levels=8, n_features_per_level=4, log2_hashmap_size=18, base_resolution=8,
                              per_level_scale=2.0)
    tables, mlp = nb.network_init_params(chess, 1)
    assert orc.Net(chess, tables, mlp, (0, 0, 0), (1, 1, 1)).levels()["n_params"] == len(tables)


def test_camera_look_at(nb):
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, 200, 100)
    du, dv, tl = np.array(cam.du), np.array(cam.dv), np.array(cam.dir_top_left)
    h = 2 * np.tan(np.radians(22.5))
    assert np.allclose(np.linalg.norm(dv), h, rtol=1e-6) and np.allclose(np.linalg.norm(du), 2 * h, rtol=1e-6)
    assert np.allclose(tl + 0.5 * du + 0.5 * dv, [0, 0, -1], atol=1e-6)
    assert dv[1] < 0  # image y goes down


def test_context_without_gpu_reports_no_device(nb):
    from nbvh_amd import hip
    if hip.device_count() > 0:
        pytest.skip("GPU present")
    with pytest.raises(nb.NbvhError) as e:
        nb.Context()
    assert e.value.status == 3


def test_cpp_driver_host_side(nb):
    """The C++ host interface (include/nbvh.hpp) driving the library: builders, .nbvh round trip,
    and the no-device error path on a machine without a GPU."""
    exe = os.path.join(ROOT, "nbvh_amd", "nbvh_driver")
    assert os.path.exists(exe), "build with `python -m nbvh_amd.build`"
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "host ok: 1280 triangles" in r.stdout
    assert "N-BVH nodes" in r.stdout and "device:" in r.stdout


def test_tcnn_parameter_order_round_trip(nb, tmp_path):
    """tiny-cuda-nn's flat parameter order (MLP matrices first, input width and output rows padded
    to 16, then the tables): the counts, the padding and the round trip through the reference's
    serialised-model file format. The order itself is tiny-cuda-nn's public convention and stays
    unpinned: no reference weights file or test exists."""
    import json
    import msgpack
    cfg = nb.network_config()
    tables, mlp = nb.network_init_params(cfg, seed=3)
    params = nb.network_to_tcnn_params(cfg, tables, mlp)
    assert len(params) == 64 * 96 + 3 * 64 * 64 + 16 * 64 + len(tables)
    # Output matrix rows 8..15 are padding (zero); its rows 0..7 are the MLP's last matrix.
    out = params[64 * 96 + 3 * 64 * 64: 64 * 96 + 3 * 64 * 64 + 16 * 64].reshape(16, 64)
    assert not out[8:].any() and np.array_equal(out[:8], mlp[-8 * 64:].reshape(8, 64))
    assert np.array_equal(params[:64 * 96].reshape(64, 96), mlp[:64 * 96].reshape(64, 96))
    assert np.array_equal(params[-len(tables):], tables)
    t2, m2 = nb.tcnn_params_to_network(cfg, params)
    assert np.array_equal(t2, tables) and np.array_equal(m2, mlp)
    # Padding is dropped on the way in.
    p2 = params.copy()
    p2[64 * 96 + 3 * 64 * 64 + 8 * 64: 64 * 96 + 3 * 64 * 64 + 16 * 64] = 7.0
    assert np.array_equal(nb.tcnn_params_to_network(cfg, p2)[1], mlp)
    with pytest.raises(nb.NbvhError):
        nb.tcnn_params_to_network(cfg, params[:-1])
    # A network whose input width is not a multiple of 16 (P*L*F = 3*2*2 = 12 -> 16).
    small = nb.network_config(n_levels=2, n_features_per_level=2, log2_hashmap_size=10)
    ts, ms = nb.network_init_params(small, seed=4)
    ps = nb.network_to_tcnn_params(small, ts, ms)
    first = ps[:64 * 16].reshape(64, 16)
    assert not first[:, 12:].any() and np.array_equal(first[:, :12], ms[:64 * 12].reshape(64, 12))
    assert np.array_equal(np.concatenate(nb.tcnn_params_to_network(small, ps)), np.concatenate([ts, ms]))
    # The serialised-model file: nlohmann::json text of Trainer::serialize (binary value dumped as
    # {"bytes": [...], "subtype": null}), and the same object as MessagePack.
    model = {"n_params": len(params), "params_type": "__half",
             "params_binary": {"bytes": list(params.astype("<f4").tobytes()), "subtype": None}}
    path = tmp_path / "model.json"
    path.write_text(json.dumps(model))
    assert np.array_equal(nb.read_tcnn_model(str(path)), params)
    mp = tmp_path / "model.msgpack"
    mp.write_bytes(msgpack.packb({"n_params": len(params), "params_binary": params.astype("<f4").tobytes()}))
    assert np.array_equal(nb.read_tcnn_model(str(mp)), params)
    model["n_params"] += 1
    path.write_text(json.dumps(model))
    with pytest.raises(nb.NbvhError):
        nb.read_tcnn_model(str(path))


def test_gather_ceiling_helper_validates_arguments():
    """The benchmark's line-rate and atomic-rate helpers load and reject invalid shapes before any
    GPU call (hipErrorInvalidValue = 1)."""
    lib = ctypes.CDLL(os.path.join(ROOT, "nbvh_amd", "libnbvh_bench.so"))
    fn = lib.nbvh_bench_gather_lines
    fn.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    ms, rate = ctypes.c_float(), ctypes.c_double()
    for args in ((64, 1, 8, 512), (1 << 20, 3, 8, 512), (1 << 20, 64, 8, 512), (1 << 20, 1, 8, 7), (1 << 20, 1, 6, 512)):
        assert fn(*args, ctypes.byref(ms), ctypes.byref(rate)) == 1, args
    add = lib.nbvh_bench_atomic_adds
    add.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
    add.restype = ctypes.c_int
    exact = ctypes.c_int32()
    for args in ((0, 8, 1), ((1 << 28) + 1, 8, 1), (1024, 6, 1), (1024, 8, 0)):
        assert add(*args, ctypes.byref(ms), ctypes.byref(rate), ctypes.byref(exact)) == 1, args
