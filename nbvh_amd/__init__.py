"""nbvh_amd -- MI355X-native N-BVH neural ray query.

Python mirror of the reference's Neural BVH Module interface
(include/cuda/neural/neural_bvh_module.h:144-241 in WeiPhil/nbvh) over the C ABI of
libnbvh_hip.so (include/nbvh.h). All compute runs in the HIP library; this layer only moves
arguments and raises NbvhError on any non-OK status.
"""
import ctypes

import numpy as np

from . import _ffi
from ._ffi import (HIT_DTYPE, NODE_DTYPE, RAY_DTYPE, NbvhError, NetworkConfig, GridConfig, QueryParams, Camera,
                   PREC_FP32, PREC_FP16, PREC_BF16, PREC_FP8, QUERY_NEURAL_CLOSEST, QUERY_NEURAL_ANY,
                   QUERY_TRIANGLE_CLOSEST, QUERY_TRIANGLE_ANY, HIT_NONE, HIT_NEURAL, HIT_TRIANGLE,
                   FLAG_STACK_OVERFLOW, FLAG_STEP_CAP, NO_NEURAL_CHILD, SEGMENT_EQUIDISTANT_EXTERIOR,
                   SEGMENT_EQUIDISTANT_INTERIOR, SEGMENT_UNIFORM, BACKGROUND_CONSTANT, BACKGROUND_SKY, TRAIN_TABLES, TRAIN_MLP,
                   TRAIN_GRAD_TABLES, TRAIN_GRAD_MLP, DEPTH_LOCAL_T, DEPTH_WORLD_T)
from ._ffi import check as _check, lib as _lib, ptr as _ptr

__all__ = ["Context", "Scene", "NeuralBVHModule", "Server", "Trainer", "Learner", "flip", "NbvhError", "network_config", "icosphere", "bvh_build",
           "nbvh_cut", "nbvh_parse", "nbvh_serialize", "geosphere", "sphere_field", "network_param_counts", "network_init_params",
           "camera_look_at", "tcnn_params_to_network", "network_to_tcnn_params", "read_tcnn_model", "leaf_segments", "hit_kind", "hit_flags", "hit_segments", "HIT_DTYPE", "NODE_DTYPE", "RAY_DTYPE"]


def hit_kind(hits):
    return hits["info"] & 0xFF


def hit_flags(hits):
    return (hits["info"] >> 8) & 0xFF


def hit_segments(hits):
    return hits["info"] >> 16


def network_config(n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                   per_level_scale=1.5, n_points=3, segment_distribution=SEGMENT_EQUIDISTANT_EXTERIOR,
                   n_neurons=64, n_hidden_layers=4, n_outputs=8, depth_mode=0, depth_linear_activation=False,
                   interpolate_depth=False, depth_interpolation_eps=0.0):
    """nbvh_network_config; defaults are the benchmark network (16 levels x 2 features, 2^19) with
    the reference's default decoding (LocalT depth, sigmoid outputs, no depth interpolation)."""
    return NetworkConfig(GridConfig(n_levels, n_features_per_level, log2_hashmap_size, base_resolution,
                                    per_level_scale), n_points, segment_distribution, n_neurons, n_hidden_layers,
                         n_outputs, depth_mode, 1 if depth_linear_activation else 0, 1 if interpolate_depth else 0,
                         depth_interpolation_eps)


# ------------------------------------------------------------------------------------------
# Host-side builders (no GPU needed)
# ------------------------------------------------------------------------------------------
def icosphere(subdivisions, seed=0, amplitude=0.1, center=(0.0, 0.0, 0.0), radius=1.0):
    """Displaced icosphere: (verts float32 [n,3], tris uint32 [m,3])."""
    L = _lib()
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    c = np.asarray(center, np.float32)
    _check(L.nbvh_mesh_icosphere(subdivisions, seed, amplitude, _ptr(c), radius, None, ctypes.byref(nv), None,
                                 ctypes.byref(nt)))
    v = np.empty((nv.value, 3), np.float32)
    t = np.empty((nt.value, 3), np.uint32)
    _check(L.nbvh_mesh_icosphere(subdivisions, seed, amplitude, _ptr(c), radius, _ptr(v), ctypes.byref(nv), _ptr(t),
                                 ctypes.byref(nt)))
    return v, t


def geosphere(frequency, seed=0, amplitude=0.1, center=(0.0, 0.0, 0.0), radius=1.0):
    """Displaced geodesic sphere of frequency k (20 k^2 triangles): (verts [n,3], tris [m,3])."""
    L = _lib()
    nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
    c = np.asarray(center, np.float32)
    _check(L.nbvh_mesh_geosphere(frequency, seed, amplitude, _ptr(c), radius, None, ctypes.byref(nv), None,
                                 ctypes.byref(nt)))
    v = np.empty((nv.value, 3), np.float32)
    t = np.empty((nt.value, 3), np.uint32)
    _check(L.nbvh_mesh_geosphere(frequency, seed, amplitude, _ptr(c), radius, _ptr(v), ctypes.byref(nv), _ptr(t),
                                 ctypes.byref(nt)))
    return v, t


def sphere_field(n_spheres=64, seed=0, freq_range=(29, 50), amplitude=0.1, radius_range=(0.04, 0.12)):
    """Procedural multi-object scene: n displaced geodesic spheres with seeded frequencies
    (20 k^2 triangles each), radii and centres in the unit cube, merged into one mesh.
    The defaults give 64 spheres of 16.8k-50k triangles, 2.1M triangles in total."""
    rng = np.random.default_rng(seed)
    vs, ts, off = [], [], 0
    for i in range(n_spheres):
        k = int(rng.integers(freq_range[0], freq_range[1] + 1))
        r = float(rng.uniform(*radius_range))
        c = rng.uniform(r, 1.0 - r, 3)
        v, t = geosphere(k, seed=int(rng.integers(1 << 62)), amplitude=amplitude, center=c, radius=r)
        vs.append(v)
        ts.append(t + np.uint32(off))
        off += len(v)
    return np.concatenate(vs), np.concatenate(ts)


def bvh_build(verts, tris, max_leaf_size=2):
    """Binned-SAH BVH2: (nodes NODE_DTYPE, prim_indices uint32)."""
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    nt = len(tris)
    nodes = np.zeros(max(2 * nt - 1, 1), NODE_DTYPE)
    prims = np.zeros(nt, np.uint32)
    n = ctypes.c_uint32()
    _check(_lib().nbvh_bvh_build(_ptr(verts), len(verts), _ptr(tris), nt, max_leaf_size, _ptr(nodes), len(nodes),
                                 ctypes.byref(n), _ptr(prims)))
    return nodes[: n.value].copy(), prims


def nbvh_cut(bvh_nodes, n_leaves):
    bvh_nodes = np.ascontiguousarray(bvh_nodes, NODE_DTYPE)
    out = np.zeros(2 * n_leaves - 1, NODE_DTYPE)
    n = ctypes.c_uint32()
    _check(_lib().nbvh_nbvh_cut(_ptr(bvh_nodes), len(bvh_nodes), n_leaves, _ptr(out), len(out), ctypes.byref(n)))
    return out[: n.value].copy()


def nbvh_parse(data):
    """Parse .nbvh bytes into N-BVH nodes (NODE_DTYPE)."""
    buf = np.frombuffer(bytes(data), np.uint8)
    n = ctypes.c_uint32()
    _check(_lib().nbvh_nbvh_parse(_ptr(buf), len(buf), None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, NODE_DTYPE)
    _check(_lib().nbvh_nbvh_parse(_ptr(buf), len(buf), _ptr(out), len(out), ctypes.byref(n)))
    return out


def nbvh_serialize(nodes):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    size = ctypes.c_size_t()
    _check(_lib().nbvh_nbvh_serialize(_ptr(nodes), len(nodes), None, 0, ctypes.byref(size)))
    out = np.zeros(size.value, np.uint8)
    _check(_lib().nbvh_nbvh_serialize(_ptr(nodes), len(nodes), _ptr(out), len(out), ctypes.byref(size)))
    return out.tobytes()


def network_param_counts(cfg):
    nt, nm = ctypes.c_uint64(), ctypes.c_uint64()
    _check(_lib().nbvh_network_param_counts(ctypes.byref(cfg), ctypes.byref(nt), ctypes.byref(nm)))
    return nt.value, nm.value


def network_init_params(cfg, seed, table_range=1e-4):
    nt, nm = network_param_counts(cfg)
    tables = np.empty(nt, np.float32)
    mlp = np.empty(nm, np.float32)
    _check(_lib().nbvh_network_init_params(ctypes.byref(cfg), seed, table_range, _ptr(tables), _ptr(mlp)))
    return tables, mlp


def tcnn_params_to_network(cfg, params):
    """tiny-cuda-nn flat parameters (the reference's serialised model) -> (tables, mlp)."""
    p = np.ascontiguousarray(params, np.float32)
    nt, nm = network_param_counts(cfg)
    tables, mlp = np.empty(nt, np.float32), np.empty(nm, np.float32)
    _check(_lib().nbvh_tcnn_params_to_network(ctypes.byref(cfg), _ptr(p), len(p), _ptr(tables), _ptr(mlp)))
    return tables, mlp


def network_to_tcnn_params(cfg, tables, mlp):
    """(tables, mlp) -> tiny-cuda-nn flat parameters (padding rows and columns zero)."""
    n = ctypes.c_uint64()
    _check(_lib().nbvh_network_to_tcnn_params(ctypes.byref(cfg), None, None, None, 0, ctypes.byref(n)))
    t = np.ascontiguousarray(tables, np.float32)
    m = np.ascontiguousarray(mlp, np.float32)
    out = np.empty(n.value, np.float32)
    _check(_lib().nbvh_network_to_tcnn_params(ctypes.byref(cfg), _ptr(t), _ptr(m), _ptr(out), n.value,
                                              ctypes.byref(n)))
    return out


def read_tcnn_model(path):
    """Flat float32 parameters of a model file written by the reference's save_network_model
    (src/cuda/neural/neural_bvh_renderer_init.cu:385-392: the JSON text of tiny-cuda-nn's
    Trainer::serialize, whose "params_binary" is a JSON binary value, dumped as
    {"bytes": [...], "subtype": ...}). A MessagePack file of the same object is read too."""
    import json
    with open(path, "rb") as f:
        raw = f.read()
    try:
        data = json.loads(raw.decode("utf-8"))
    except (UnicodeDecodeError, ValueError):
        import msgpack
        data = msgpack.unpackb(raw, raw=False)
    blob = data["params_binary"]
    if isinstance(blob, dict):
        blob = bytes(blob["bytes"])
    params = np.frombuffer(bytes(blob), "<f4").copy()
    n = int(data.get("n_params", len(params)))
    if len(params) != n:
        raise NbvhError(_ffi.ERR_FORMAT, f"params_binary holds {len(params)} floats, n_params says {n}")
    return params


def camera_look_at(eye, target, up, fovy_degrees, width, height):
    cam = Camera()
    e, t, u = (np.asarray(x, np.float32) for x in (eye, target, up))
    _check(_lib().nbvh_camera_look_at(_ptr(e), _ptr(t), _ptr(u), fovy_degrees, width, height, ctypes.byref(cam)))
    return cam


def camera_array(cam):
    """Camera as float32[12] (origin, du, dv, dir_top_left)."""
    return np.array(list(cam.origin) + list(cam.du) + list(cam.dv) + list(cam.dir_top_left), np.float32)


# ------------------------------------------------------------------------------------------
# Device objects
# ------------------------------------------------------------------------------------------
class Context:
    """nbvh_create / nbvh_destroy: one GPU, an HBM budget, counters for concurrent queries."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        info = _ffi.CreateInfo(device, 0, hbm_budget_bytes)
        h = ctypes.c_void_p()
        _check(_lib().nbvh_create(ctypes.byref(info), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _lib().nbvh_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, status):
        _check(status, self.handle)

    def memory_usage(self):
        used, budget = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(_lib().nbvh_memory_usage(self.handle, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    def raygen(self, cam, width, height, d_rays_ptr, sample=0, frame=0, stream=None):
        """Primary rays into device memory at d_rays_ptr (width*height nbvh_ray records)."""
        self.check(_lib().nbvh_raygen(self.handle, ctypes.byref(cam), width, height, sample, frame,
                                      ctypes.c_void_p(d_rays_ptr), ctypes.c_void_p(stream or 0)))


class Scene:
    """A device-resident scene: triangle BVH, N-BVH, encoding box and network weights."""

    def __init__(self, ctx, verts=None, tris=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if verts is None:
            ctx.check(_lib().nbvh_scene_create_empty(ctx.handle, ctypes.byref(h)))
        else:
            v = np.ascontiguousarray(verts, np.float32)
            t = np.ascontiguousarray(tris, np.uint32)
            ctx.check(_lib().nbvh_scene_create(ctx.handle, _ptr(v), len(v), _ptr(t), len(t), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _lib().nbvh_scene_destroy(self.handle)
            self.handle = None

    def bvh(self):
        nn, npr = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx.check(_lib().nbvh_scene_get_bvh(self.handle, None, 0, ctypes.byref(nn), None, 0, ctypes.byref(npr)))
        nodes = np.zeros(nn.value, NODE_DTYPE)
        prims = np.zeros(npr.value, np.uint32)
        self.ctx.check(_lib().nbvh_scene_get_bvh(self.handle, _ptr(nodes), len(nodes), ctypes.byref(nn), _ptr(prims),
                                                 len(prims), ctypes.byref(npr)))
        return nodes, prims

    # N-BVH (load_neural_bvh, neural_bvh_renderer_accel.cpp:208)
    def build_nbvh(self, n_leaves):
        self.ctx.check(_lib().nbvh_scene_build_nbvh(self.handle, n_leaves))

    def load_neural_bvh(self, path_or_bytes):
        data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
        buf = np.frombuffer(bytes(data), np.uint8)
        self.ctx.check(_lib().nbvh_scene_load_nbvh(self.handle, _ptr(buf), len(buf)))

    def set_nbvh(self, nodes):
        nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
        self.ctx.check(_lib().nbvh_scene_set_nbvh(self.handle, _ptr(nodes), len(nodes)))

    def nbvh(self):
        n = ctypes.c_uint32()
        self.ctx.check(_lib().nbvh_scene_get_nbvh(self.handle, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, NODE_DTYPE)
        self.ctx.check(_lib().nbvh_scene_get_nbvh(self.handle, _ptr(out), len(out), ctypes.byref(n)))
        return out

    def set_bbox(self, bmin, bmax):
        a, b = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
        self.ctx.check(_lib().nbvh_scene_set_bbox(self.handle, _ptr(a), _ptr(b)))

    def bbox(self):
        a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
        self.ctx.check(_lib().nbvh_scene_get_bbox(self.handle, _ptr(a), _ptr(b)))
        return a, b

    # Network (load_network_model, neural_bvh_renderer_init.cu:394)
    def load_network_model(self, cfg, tables, mlp):
        t = np.ascontiguousarray(tables, np.float32)
        m = np.ascontiguousarray(mlp, np.float32)
        self.ctx.check(_lib().nbvh_scene_load_network(self.handle, ctypes.byref(cfg), _ptr(t), len(t), _ptr(m),
                                                      len(m)))

    def load_tcnn_model(self, cfg, path_or_params):
        """load_network_model from the reference's serialised model (a file path) or from a flat
        tiny-cuda-nn parameter array."""
        params = read_tcnn_model(path_or_params) if isinstance(path_or_params, (str, bytes)) or hasattr(
            path_or_params, "__fspath__") else path_or_params
        tables, mlp = tcnn_params_to_network(cfg, params)
        self.load_network_model(cfg, tables, mlp)

    def init_network(self, cfg, seed, table_range=1e-4):
        self.ctx.check(_lib().nbvh_scene_init_network(self.handle, ctypes.byref(cfg), seed, table_range))

    def network(self):
        cfg = NetworkConfig()
        self.ctx.check(_lib().nbvh_scene_get_network(self.handle, ctypes.byref(cfg), None, 0, None, 0))
        nt, nm = network_param_counts(cfg)
        t = np.zeros(nt, np.float32)
        m = np.zeros(nm, np.float32)
        self.ctx.check(_lib().nbvh_scene_get_network(self.handle, ctypes.byref(cfg), _ptr(t), nt, _ptr(m), nm))
        return cfg, t, m


class NeuralBVHModule:
    """Inference side of the reference's NeuralBVHModule / NeuralPathModule.

    neural_bvh_traversal runs the whole per-ray leaf loop of
    NeuralPathModule::neural_bvh_traversal (src/cuda/neural/path/path_module.cu:54-137) and returns
    one hit record per ray; network_inference replaces tcnn::Network::inference (path_module.cu:92).
    """

    def __init__(self, ctx, scene):
        self.ctx = ctx
        self.scene = scene

    @staticmethod
    def params(kind=QUERY_NEURAL_CLOSEST, precision=PREC_FP32, lod=0, max_traversal_steps=0,
               visibility_threshold=0.0, image_width=0, sort_rays=False, wavefront=False, array_order=False,
               segment_seed=0):
        """nbvh_query_params. sort_rays: trace in the coherence order; array_order: in array order
        even where the scene's launch shape (ray_order, set by autotune) would sort; segment_seed:
        the rays' LCG streams under the uniform segment distribution."""
        flags = ((_ffi.QUERY_SORT_RAYS if sort_rays else 0) | (_ffi.QUERY_WAVEFRONT if wavefront else 0) |
                 (_ffi.QUERY_ARRAY_ORDER if array_order else 0))
        return QueryParams(kind, precision, lod, max_traversal_steps, visibility_threshold, image_width, flags,
                           segment_seed)

    def neural_bvh_traversal(self, rays, kind=QUERY_NEURAL_CLOSEST, precision=PREC_FP32, lod=0,
                             max_traversal_steps=0, visibility_threshold=0.0, image_width=0, sort_rays=False,
                             wavefront=False, array_order=False, segment_seed=0):
        """Host rays (RAY_DTYPE) -> host hits (HIT_DTYPE); synchronous."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        p = self.params(kind, precision, lod, max_traversal_steps, visibility_threshold, image_width, sort_rays,
                        wavefront, array_order, segment_seed)
        self.ctx.check(_lib().nbvh_query_host(self.ctx.handle, self.scene.handle, ctypes.byref(p), _ptr(rays),
                                              len(rays), _ptr(hits)))
        return hits

    def query_device(self, params, d_rays_ptr, n, d_hits_ptr, stream=None):
        """Asynchronous query on device buffers (raw device pointers), on `stream` (hipStream_t)."""
        self.ctx.check(_lib().nbvh_query(self.ctx.handle, self.scene.handle, ctypes.byref(params),
                                         ctypes.c_void_p(d_rays_ptr), n, ctypes.c_void_p(d_hits_ptr),
                                         ctypes.c_void_p(stream or 0)))

    def query_device_noalloc(self, params, d_rays_ptr, n, d_hits_ptr, stream=None):
        """query_device through nbvh_query_noalloc: no allocation, scratch in the library's static
        per-GPU array (calls overlapping on one GPU race; sorted queries up to 524,288 rays)."""
        self.ctx.check(_lib().nbvh_query_noalloc(self.ctx.handle, self.scene.handle, ctypes.byref(params),
                                                 ctypes.c_void_p(d_rays_ptr), n, ctypes.c_void_p(d_hits_ptr),
                                                 ctypes.c_void_p(stream or 0)))

    def autotune(self, params, d_rays_ptr, n, d_hits_ptr, stream=None):
        """Time the candidate launch shapes (and, for batches without an image order, the
        coherence order) on these device rays; keep the fastest for the scene and precision.
        Returns ((blocks_per_cu, queue_bands, ray_order), ms)."""
        best, ms = _ffi.LaunchShape(), ctypes.c_float()
        self.ctx.check(_lib().nbvh_autotune(self.ctx.handle, self.scene.handle, ctypes.byref(params),
                                            ctypes.c_void_p(d_rays_ptr), n, ctypes.c_void_p(d_hits_ptr),
                                            ctypes.c_void_p(stream or 0), ctypes.byref(best), ctypes.byref(ms)))
        return (best.blocks_per_cu, best.queue_bands, best.ray_order), ms.value

    def set_launch_shape(self, precision, blocks_per_cu=0, queue_bands=0, ray_order=0):
        shape = _ffi.LaunchShape(blocks_per_cu, queue_bands, ray_order)
        self.ctx.check(_lib().nbvh_scene_set_launch_shape(self.scene.handle, precision, ctypes.byref(shape)))

    def launch_shape(self, precision):
        """(blocks_per_cu, queue_bands, ray_order) of the scene for this precision."""
        shape = _ffi.LaunchShape()
        self.ctx.check(_lib().nbvh_scene_get_launch_shape(self.scene.handle, precision, ctypes.byref(shape)))
        return shape.blocks_per_cu, shape.queue_bands, shape.ray_order

    def segment_labels(self, rays, mint, maxt):
        """Ground-truth targets [n, 8] of host segments (rays clipped to [mint, maxt)) from the
        scene's triangles (vis, LocalT depth, (n+1)/2 normal, albedo); synchronous."""
        from . import hip
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        if n == 0:
            return np.zeros((0, 8), np.float32)
        dr, d0, d1 = hip.Buffer.from_array(rays), hip.Buffer.from_array(np.asarray(mint, np.float32)), \
            hip.Buffer.from_array(np.asarray(maxt, np.float32))
        dt = hip.Buffer(n * 32)
        self.ctx.check(_lib().nbvh_segment_labels(self.ctx.handle, self.scene.handle, ctypes.c_void_p(dr.ptr),
                                                  ctypes.c_void_p(d0.ptr), ctypes.c_void_p(d1.ptr), n,
                                                  ctypes.c_void_p(dt.ptr), None))
        hip.synchronize()
        return dt.download(np.float32).reshape(n, 8)

    def scatter_diffuse_device(self, d_rays_ptr, d_hits_ptr, n, d_next_ptr, sample=0, bounce=0, stream=None):
        """Next rays of a diffuse bounce (nbvh_scatter_diffuse); rays without a hit get tmax = -1."""
        self.ctx.check(_lib().nbvh_scatter_diffuse(self.ctx.handle, self.scene.handle, ctypes.c_void_p(d_rays_ptr),
                                                   ctypes.c_void_p(d_hits_ptr), n, sample, bounce,
                                                   ctypes.c_void_p(d_next_ptr), ctypes.c_void_p(stream or 0)))

    def query_hybrid_device(self, triangles, params, d_rays_ptr, n, d_hits_ptr, stream=None):
        """Closest (or any) hit over this module's neural scene and the triangle scene `triangles`."""
        self.ctx.check(_lib().nbvh_query_hybrid(self.ctx.handle, self.scene.handle, triangles.handle,
                                                ctypes.byref(params), ctypes.c_void_p(d_rays_ptr), n,
                                                ctypes.c_void_p(d_hits_ptr), ctypes.c_void_p(stream or 0)))

    def query_hybrid(self, triangles, rays, kind=QUERY_NEURAL_CLOSEST, precision=PREC_FP32, **kw):
        """Host rays -> host hits of a hybrid query (synchronous)."""
        from . import hip
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        if len(rays) == 0:
            return np.zeros(0, HIT_DTYPE)
        dr, dh = hip.Buffer.from_array(rays), hip.Buffer(len(rays) * 48)
        self.query_hybrid_device(triangles, self.params(kind, precision, **kw), dr.ptr, len(rays), dh.ptr)
        hip.synchronize()
        return dh.download(HIT_DTYPE)

    def path_params(self, width, height, precision=PREC_FP32, max_depth=5, rr_start_depth=0, sample=0,
                    background=BACKGROUND_SKY, background_color=(0.0, 0.0, 0.0), environment_albedo=(0.8, 0.8, 0.8),
                    normal_offset=0.0, sort_bounces=True, lod=0, bounce_lod=None, max_traversal_steps=0,
                    visibility_threshold=0.0, secondary_visibility_threshold=0.0, triangles=False,
                    wavefront_bounces=False, workspace=None):
        """nbvh_path_params of a frame (see path_trace_device). workspace: a device buffer of at
        least path_trace_workspace_size bytes (hip.Buffer), or None."""
        p = _ffi.PathParams()
        p.query = self.params(QUERY_TRIANGLE_CLOSEST if triangles else QUERY_NEURAL_CLOSEST, precision, lod,
                              max_traversal_steps, visibility_threshold,
                              image_width=width if width % 8 == 0 and height % 8 == 0 else 0)
        p.secondary_visibility_threshold = secondary_visibility_threshold
        p.bounce_query_flags = ((_ffi.QUERY_SORT_RAYS if sort_bounces else 0) |
                                (_ffi.QUERY_WAVEFRONT if wavefront_bounces else 0))
        p.max_depth, p.rr_start_depth, p.sample, p.background = max_depth, rr_start_depth, sample, background
        p.background_color[:] = [float(c) for c in background_color]
        p.environment_albedo[:] = [float(c) for c in environment_albedo]
        p.normal_offset = normal_offset
        if bounce_lod is not None:
            p.path_lod_switch, p.lod_switch = 1, bounce_lod
        if workspace is not None:
            p.workspace, p.workspace_bytes = workspace.ptr, workspace.nbytes
        return p

    def path_trace_workspace_size(self, params, width, height, environment=None):
        """Device bytes nbvh_path_trace needs as its workspace for these parameters."""
        b = ctypes.c_uint64()
        self.ctx.check(_lib().nbvh_path_trace_workspace_size(self.ctx.handle, self.scene.handle,
                                                             environment.handle if environment is not None else None,
                                                             ctypes.byref(params), width, height, ctypes.byref(b)))
        return int(b.value)

    def path_trace_params_device(self, params, camera, width, height, d_accum_ptr, environment=None, stream=None,
                                 stats=True):
        """nbvh_path_trace with ready-made params; returns the rays traced per depth (stats=True) or
        None (stats=False: no read-back, as under graph capture)."""
        st = _ffi.PathStats() if stats else None
        self.ctx.check(_lib().nbvh_path_trace(self.ctx.handle, self.scene.handle,
                                              environment.handle if environment is not None else None,
                                              ctypes.byref(params), ctypes.byref(camera), width, height,
                                              ctypes.c_void_p(d_accum_ptr), ctypes.byref(st) if st is not None else None,
                                              ctypes.c_void_p(stream or 0)))
        return [int(r) for r in st.rays[:params.max_depth]] if st is not None else None

    def path_trace_device(self, camera, width, height, d_accum_ptr, environment=None, stream=None, **kw):
        """One frame of 1-spp diffuse path tracing into the device float4 buffer d_accum_ptr
        (width*height; rgb sum and sample count). `environment`: a triangle Scene traced with the
        neural object (hybrid), or None. Keywords: those of path_params -- bounce_lod: N-BVH LoD of
        the rays at depth > 1 (None: lod); secondary_visibility_threshold: threshold of the rays at
        depth > 1 (0: visibility_threshold); triangles=True renders the object's own triangles with
        environment_albedo instead of its N-BVH and network (the reference's ignore_neural_blas).
        Returns the rays traced per depth."""
        p = self.path_params(width, height, **kw)
        return self.path_trace_params_device(p, camera, width, height, d_accum_ptr, environment, stream)

    def path_trace(self, camera, width, height, samples=1, **kw):
        """Host image [height, width, 4] (rgb sums, sample count) of `samples` frames."""
        from . import hip
        accum = hip.Buffer(max(1, width * height) * 16)
        rays = []
        for s in range(samples):
            rays.append(self.path_trace_device(camera, width, height, accum.ptr, sample=s, **kw))
        hip.synchronize()
        img = accum.download(np.float32, width * height * 4).reshape(height, width, 4)
        return img, rays

    @staticmethod
    def _parts(parts):
        arr = (_ffi.QueryPart * len(parts))()
        for i, (r, n, h, any_hit) in enumerate(parts):
            arr[i] = _ffi.QueryPart(r, h, n, 1 if any_hit else 0)
        return arr

    def query_parts_device(self, params, parts, stream=None):
        """One launch over several device ray arrays: parts = [(d_rays_ptr, n, d_hits_ptr, any_hit)].
        params.kind selects the family (neural or triangle); any_hit is per part."""
        self.ctx.check(_lib().nbvh_query_parts(self.ctx.handle, self.scene.handle, ctypes.byref(params),
                                               self._parts(parts), len(parts), ctypes.c_void_p(stream or 0)))

    def autotune_parts(self, params, parts, stream=None):
        """autotune over a multi-part launch (parts as for query_parts_device).
        Returns ((blocks_per_cu, queue_bands, ray_order), ms)."""
        best, ms = _ffi.LaunchShape(), ctypes.c_float()
        self.ctx.check(_lib().nbvh_autotune_parts(self.ctx.handle, self.scene.handle, ctypes.byref(params),
                                                  self._parts(parts), len(parts), ctypes.c_void_p(stream or 0),
                                                  ctypes.byref(best), ctypes.byref(ms)))
        return (best.blocks_per_cu, best.queue_bands, best.ray_order), ms.value

    def network_inference(self, inputs, precision=PREC_FP32):
        """inputs float32 [n, 3P] -> (outputs [n, 8] post-sigmoid, logits [n, 8])."""
        x = np.ascontiguousarray(inputs, np.float32)
        n = x.shape[0]
        out = np.zeros((n, 8), np.float32)
        logits = np.zeros((n, 8), np.float32)
        self.ctx.check(_lib().nbvh_network_inference_host(self.ctx.handle, self.scene.handle, precision, _ptr(x), n,
                                                          _ptr(out), _ptr(logits)))
        return out, logits

    def network_inference_device(self, precision, d_in_ptr, n, d_out_ptr, d_logits_ptr=0, stream=None):
        self.ctx.check(_lib().nbvh_network_inference(self.ctx.handle, self.scene.handle, precision,
                                                     ctypes.c_void_p(d_in_ptr), n, ctypes.c_void_p(d_out_ptr),
                                                     ctypes.c_void_p(d_logits_ptr or 0), ctypes.c_void_p(stream or 0)))


def leaf_segments(nbvh, n, seed=0, lod=0):
    """Training segments in the spirit of the reference's leaf sampling: a uniformly chosen N-BVH
    leaf (at LoD `lod`), a uniform point in its box and a uniform direction; the segment is the
    line's passage through the box, as distances from an origin 3 units back along the line.
    Returns (rays, mint, maxt): encode with the network's segment encoding and label with
    NeuralBVHModule.segment_labels. (The reference's granular, loss-driven node sampling and its
    tree-cut learning are not restated.)"""
    rng = np.random.default_rng(seed)
    nbvh = np.asarray(nbvh, NODE_DTYPE)
    pc = nbvh["prim_count"]
    leaf = ((pc > 0) & (pc <= 2 + lod)) | (nbvh["left_or_prim_offset"] == NO_NEURAL_CHILD)
    leaves = nbvh[leaf]
    b = leaves[rng.integers(0, len(leaves), n)]
    blo, bhi = b["aabb_min"].astype(np.float64), b["aabb_max"].astype(np.float64)
    p = rng.uniform(blo, bhi)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = p - 3.0 * d
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (blo - o) / d, (bhi - o) / d
    mint, maxt = np.max(np.minimum(t0, t1), axis=1), np.min(np.maximum(t0, t1), axis=1)
    rays = np.zeros(n, RAY_DTYPE)
    rays["o"], rays["d"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 1e-4, np.float32(3.4028234663852886e38)
    return rays, mint.astype(np.float32), maxt.astype(np.float32)


class Trainer:
    """Training of a scene's network (nbvh_trainer_*): the replacement of the reference's
    trainer->training_step on encoded segments and their 8 targets, composite loss and Adam with
    the reference's defaults, including its exponential learning-rate decay (x decay_base after
    update decay_start and every decay_interval updates after it; decay_interval=0 turns it off).
    commit() writes the trained network into the scene."""

    def __init__(self, ctx, scene, learning_rate=1e-2, beta1=0.9, beta2=0.999, epsilon=1e-8, l2_reg=1e-8,
                 max_batch=1 << 18, decay_start=10000, decay_interval=2000, decay_base=0.33):
        self.ctx, self.scene = ctx, scene
        cfg = _ffi.TrainConfig(learning_rate, beta1, beta2, epsilon, l2_reg, max_batch, decay_start,
                               decay_interval, decay_base)
        self.handle = ctypes.c_void_p()
        self.ctx.check(_lib().nbvh_trainer_create(ctx.handle, scene.handle, ctypes.byref(cfg),
                                                  ctypes.byref(self.handle)))
        self.max_batch = max_batch

    def step_device(self, d_inputs_ptr, d_targets_ptr, n, gradients_only=False, stream=None, want_loss=True):
        """One step on device buffers (n x 9 inputs, n x 8 targets). want_loss: wait for the step
        and return its loss; otherwise the step is only queued on the stream (returns None), so
        steps can be issued back to back."""
        loss = ctypes.c_float()
        self.ctx.check(_lib().nbvh_trainer_step(self.handle, ctypes.c_void_p(d_inputs_ptr),
                                                ctypes.c_void_p(d_targets_ptr), n,
                                                _ffi.TRAIN_GRADIENTS_ONLY if gradients_only else 0,
                                                ctypes.byref(loss) if want_loss else None,
                                                ctypes.c_void_p(stream or 0)))
        return loss.value if want_loss else None

    def step(self, inputs, targets, gradients_only=False):
        from . import hip
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        if len(x) == 0:
            return self.step_device(0, 0, 0, gradients_only)
        dx, dt = hip.Buffer.from_array(x), hip.Buffer.from_array(t)
        return self.step_device(dx.ptr, dt.ptr, len(x), gradients_only)

    def get(self, what):
        """One of the trainer's arrays: _ffi.TRAIN_TABLES / TRAIN_MLP / TRAIN_GRAD_TABLES / TRAIN_GRAD_MLP."""
        n = ctypes.c_uint64()
        self.ctx.check(_lib().nbvh_trainer_get(self.handle, what, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.float32)
        self.ctx.check(_lib().nbvh_trainer_get(self.handle, what, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def commit(self):
        self.ctx.check(_lib().nbvh_trainer_commit(self.handle))

    def close(self):
        if self.handle:
            _lib().nbvh_trainer_destroy(self.handle)
            self.handle = ctypes.c_void_p()


def flip(ctx, reference, test, ppd=0.0):
    """LDR-FLIP of two host accumulation images [h, w, 4] (rgb sum, sample count), the reference's
    validation metric: (mean error, error map [h, w]). ppd 0: the reference's viewing geometry."""
    from . import hip
    ref = np.ascontiguousarray(reference, np.float32)
    tst = np.ascontiguousarray(test, np.float32)
    if ref.shape != tst.shape or ref.ndim != 3 or ref.shape[2] != 4:
        raise ValueError("flip takes two [h, w, 4] accumulation images of one size")
    h, w = ref.shape[:2]
    dr, dt, dm = hip.Buffer.from_array(ref), hip.Buffer.from_array(tst), hip.Buffer(h * w * 4)
    mean = ctypes.c_float()
    ctx.check(_lib().nbvh_flip(ctx.handle, ctypes.c_void_p(dr.ptr), ctypes.c_void_p(dt.ptr), w, h, ppd,
                               ctypes.c_void_p(dm.ptr), ctypes.byref(mean), None))
    return mean.value, dm.download(np.float32, h * w).reshape(h, w)


class Learner:
    """Tree-cut learning of a scene's N-BVH (nbvh_learner_*), the reference's training loop
    (src/cuda/neural/neural_bvh_renderer.cu:141-297): each step samples training segments from the
    N-BVH being learned, trains the network on them, adds every sample's loss to its node and splits
    the leaves of largest loss heuristic on the split schedule. Creating it replaces the scene's
    N-BVH by the cut of its triangle BVH at neural_bvh_min_depth. Keyword arguments override the
    reference's defaults (nbvh_learner_config field names)."""

    def __init__(self, ctx, scene, **overrides):
        self.ctx, self.scene = ctx, scene
        cfg = _ffi.LearnerConfig()
        _check(_lib().nbvh_learner_config_defaults(ctypes.byref(cfg)))
        for k, v in overrides.items():
            if not hasattr(cfg, k):
                raise TypeError(f"unknown learner option {k}")
            setattr(cfg, k, v)
        self.config = cfg
        self.handle = ctypes.c_void_p()
        self.ctx.check(_lib().nbvh_learner_create(ctx.handle, scene.handle, ctypes.byref(cfg),
                                                  ctypes.byref(self.handle)))

    @staticmethod
    def defaults():
        cfg = _ffi.LearnerConfig()
        _check(_lib().nbvh_learner_config_defaults(ctypes.byref(cfg)))
        return {name: getattr(cfg, name) for name, _ in cfg._fields_ if name != "reserved"}

    def sample(self, step, stream=None):
        """Sample the batch of training step `step` into the learner's buffers."""
        self.ctx.check(_lib().nbvh_learner_sample(self.handle, step, ctypes.c_void_p(stream or 0)))

    def update_device(self, d_sample_loss_ptr, stream=None):
        """Tree-cut update of the last sampled batch from per-slot losses (device floats)."""
        self.ctx.check(_lib().nbvh_learner_update(self.handle, ctypes.c_void_p(d_sample_loss_ptr),
                                                  ctypes.c_void_p(stream or 0)))

    def update(self, sample_loss):
        from . import hip
        d = hip.Buffer.from_array(np.ascontiguousarray(sample_loss, np.float32))
        self.update_device(d.ptr)

    def step(self, trainer, stream=None):
        """One whole step (sample, train, tree-cut update); returns the training loss."""
        loss = ctypes.c_float()
        self.ctx.check(_lib().nbvh_learner_step(self.handle, trainer.handle, ctypes.byref(loss),
                                                ctypes.c_void_p(stream or 0)))
        return loss.value

    def state(self):
        st = _ffi.LearnerState()
        self.ctx.check(_lib().nbvh_learner_get_state(self.handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_ if name != "reserved"}

    def get(self, what):
        """A learner array: _ffi.LEARN_INPUTS / TARGETS (float, reshaped per slot), NODES (uint32),
        TREE_CUT (TREE_CUT_DTYPE), SPLITS (uint32), NBVH (NODE_DTYPE)."""
        dtype = {_ffi.LEARN_INPUTS: np.float32, _ffi.LEARN_TARGETS: np.float32, _ffi.LEARN_NODES: np.uint32,
                 _ffi.LEARN_TREE_CUT: _ffi.TREE_CUT_DTYPE, _ffi.LEARN_SPLITS: np.uint32,
                 _ffi.LEARN_NBVH: NODE_DTYPE}[what]
        n = ctypes.c_uint64()
        self.ctx.check(_lib().nbvh_learner_get(self.handle, what, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype)
        self.ctx.check(_lib().nbvh_learner_get(self.handle, what, _ptr(out), n.value, ctypes.byref(n)))
        if what == _ffi.LEARN_INPUTS:
            return out.reshape(self.config.batch_size, -1)
        if what == _ffi.LEARN_TARGETS:
            return out.reshape(self.config.batch_size, 8)
        return out

    def close(self):
        if self.handle:
            _lib().nbvh_learner_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class Server:
    """Continuous-batching request loop (nbvh_server_*): any thread submits ray batches against
    resident scenes and waits on tickets; queued requests with the same scene and parameters share
    launches, up to n_streams launches overlap on the GPU, host-buffer requests are admitted by
    the staging arena's capacity."""

    def __init__(self, ctx, n_streams=4, max_batch_rays=1 << 21, staging_bytes=256 << 20):
        self.ctx = ctx
        cfg = _ffi.ServerConfig(n_streams, max_batch_rays, staging_bytes)
        h = ctypes.c_void_p()
        ctx.check(_lib().nbvh_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._keep = {}  # ticket -> host arrays kept alive until the request is collected

    def close(self):
        if self.handle:
            _lib().nbvh_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, scene, params, rays, hits=None):
        """Host request: rays (RAY_DTYPE array) -> a ticket; the hits array (allocated here unless
        given) is filled when the request completes and returned by wait()."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        if hits is None:
            hits = np.zeros(len(rays), HIT_DTYPE)
        t = ctypes.c_uint64()
        self.ctx.check(_lib().nbvh_request_submit(self.handle, scene.handle, ctypes.byref(params), _ptr(rays),
                                                  len(rays), _ptr(hits), 0, ctypes.byref(t)))
        self._keep[t.value] = (rays, hits)
        return t.value

    def submit_device(self, scene, params, d_rays_ptr, n, d_hits_ptr):
        """Device request (zero-copy): n rays at d_rays_ptr, hit records to d_hits_ptr."""
        t = ctypes.c_uint64()
        self.ctx.check(_lib().nbvh_request_submit(self.handle, scene.handle, ctypes.byref(params),
                                                  ctypes.c_void_p(d_rays_ptr), n, ctypes.c_void_p(d_hits_ptr),
                                                  _ffi.REQUEST_DEVICE_BUFFERS, ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout_us=-1):
        """Block until the request completes -> (hits or None for device requests, info dict).
        Raises NbvhError (status 8) on timeout, leaving the ticket valid."""
        info = _ffi.RequestInfo()
        st = _lib().nbvh_request_wait(self.handle, ticket, timeout_us, ctypes.byref(info))
        if st == _ffi.ERR_TIMEOUT:
            raise NbvhError(st, "request not complete")
        kept = self._keep.pop(ticket, None)
        self.ctx.check(st)
        d = {k: getattr(info, k) for k in ("status", "n_launches", "batch_parts", "submit_ns", "launch_ns",
                                           "complete_ns")}
        return (kept[1] if kept else None), d

    def stats(self):
        st = _ffi.ServerStats()
        self.ctx.check(_lib().nbvh_server_get_stats(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}
