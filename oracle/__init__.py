"""CPU oracle of the N-BVH neural ray query -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's algorithm (oracle/nbvh_oracle.cpp, citing the reference file
and line of every function). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may use it, and only as the checker / CPU baseline; the product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("aabb_min", "<f4", 3), ("prim_count", "<u4"), ("aabb_max", "<f4", 3),
                       ("left_or_prim_offset", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<u4"), ("node_id", "<u4"), ("info", "<u4"),
                      ("normal_or_uv", "<f4", 3), ("logit", "<f4"), ("albedo", "<f4", 3), ("reserved", "<f4")])
SEG_DTYPE = np.dtype([("ray", "<u4"), ("node", "<u4"), ("mint", "<f4"), ("maxt", "<f4"), ("logit", "<f4", 8)])

FP32, FP64 = 0, 1


class OracleNet(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_uint32), ("n_features", ctypes.c_uint32),
                ("log2_hashmap_size", ctypes.c_uint32), ("base_resolution", ctypes.c_uint32),
                ("per_level_scale", ctypes.c_float), ("n_points", ctypes.c_uint32),
                ("distribution", ctypes.c_uint32), ("n_neurons", ctypes.c_uint32),
                ("n_hidden_layers", ctypes.c_uint32), ("n_outputs", ctypes.c_uint32),
                ("tables", ctypes.c_void_p), ("mlp", ctypes.c_void_p),
                ("bbox_min", ctypes.c_float * 3), ("bbox_max", ctypes.c_float * 3),
                ("depth_mode", ctypes.c_uint32), ("depth_linear", ctypes.c_uint32),
                ("interpolate_depth", ctypes.c_uint32), ("interp_eps", ctypes.c_float),
                ("segment_seed", ctypes.c_uint32)]


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
        sig = {
            "oracle_murmur_mix": ([u32, u32], u32), "oracle_murmur_fmix": ([u32], u32),
            "oracle_lcg_seed": ([u32, u32, u32], u32), "oracle_lcg_next": ([vp], u32),
            "oracle_lcg_float": ([vp], f32), "oracle_set_num_threads": ([ctypes.c_int], None), "oracle_expf": ([f32], f32), "oracle_sigmoidf": ([f32], f32),
            "oracle_num_threads": ([], ctypes.c_int),
            "oracle_raygen": ([vp, u32, u32, u32, u32, vp], None),
            "oracle_scatter_diffuse": ([vp, vp, vp, vp, u32, u32, u32, vp], None),
            "oracle_path_shade": ([vp, vp, u32, u32, u32, u32, vp, vp, f32, u32, vp, vp, vp, u32, vp, vp, vp], u32),
            "oracle_slab": ([vp, vp, vp, vp, vp], ctypes.c_int),
            "oracle_dilate": ([vp, f32, vp, vp], None),
            "oracle_nbvh_leaf_sequence": ([vp, u32, vp, f32, vp, vp, vp, u32], ctypes.c_int),
            "oracle_grid_levels": ([vp, vp, vp, vp, vp, vp], u64),
            "oracle_encode": ([vp, vp, vp, vp, u32, u32, vp], None),
            "oracle_grid": ([vp, vp, u32, u32, vp], None),
            "oracle_network": ([vp, vp, u32, u32, vp, vp], None),
            "oracle_neural_query": ([vp, vp, vp, u32, u32, u32, u32, u32, f32, vp, vp, u64, vp], None),
            "oracle_tri_query": ([vp, vp, vp, vp, vp, u32, u32, vp], None),
            "oracle_tri_brute": ([vp, vp, u32, vp, u32, vp], None),
            "oracle_learn_sample": ([vp] * 15, None),
            "oracle_sincosf": ([vp, u32, vp, vp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Net:
    """Network + encoding box for the oracle (keeps the arrays alive)."""

    def __init__(self, cfg, tables, mlp, bbox_min, bbox_max):
        self.tables = np.ascontiguousarray(tables, np.float32)
        self.mlp = np.ascontiguousarray(mlp, np.float32)
        g = cfg.grid
        self.s = OracleNet(g.n_levels, g.n_features_per_level, g.log2_hashmap_size, g.base_resolution,
                           g.per_level_scale, cfg.n_points, cfg.segment_distribution, cfg.n_neurons,
                           cfg.n_hidden_layers, cfg.n_outputs, self.tables.ctypes.data, self.mlp.ctypes.data,
                           (ctypes.c_float * 3)(*bbox_min), (ctypes.c_float * 3)(*bbox_max),
                           getattr(cfg, "depth_mode", 0), getattr(cfg, "depth_linear_activation", 0),
                           getattr(cfg, "interpolate_depth", 0),
                           getattr(cfg, "depth_interpolation_eps", 0.0) or 1e-3)
        self.cfg = cfg

    @property
    def n_points(self):
        return self.s.n_points

    def levels(self):
        L = self.s.n_levels
        scale = np.zeros(L, np.float32)
        res, size, off, dense = (np.zeros(L, np.uint32) for _ in range(4))
        total = lib().oracle_grid_levels(ctypes.byref(self.s), _p(scale), _p(res), _p(size), _p(off), _p(dense))
        return dict(scale=scale, res=res, size=size, offset=off, dense=dense, n_params=total)

    def encode(self, rays, mint, maxt, mode=FP32, segment_seed=0):
        """Encoded points of segments (rays[i], mint[i], maxt[i]); under the uniform distribution
        segment i takes the first n_points draws of stream lcg_seed(segment_seed, 0, i)."""
        self.s.segment_seed = segment_seed
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        mint = np.ascontiguousarray(mint, np.float32)
        maxt = np.ascontiguousarray(maxt, np.float32)
        out = np.zeros((len(rays), 3 * self.n_points), np.float32 if mode == FP32 else np.float64)
        lib().oracle_encode(ctypes.byref(self.s), _p(rays), _p(mint), _p(maxt), len(rays), mode, _p(out))
        return out

    def grid(self, x, mode=FP32):
        x = np.ascontiguousarray(x, np.float32 if mode == FP32 else np.float64).reshape(-1, 3)
        LF = self.s.n_levels * self.s.n_features
        out = np.zeros((len(x), LF), x.dtype)
        lib().oracle_grid(ctypes.byref(self.s), _p(x), len(x), mode, _p(out))
        return out

    def network(self, inputs, mode=FP64):
        x = np.ascontiguousarray(inputs, np.float32)
        n = x.shape[0]
        logits = np.zeros((n, self.s.n_outputs), np.float64)
        outs = np.zeros((n, self.s.n_outputs), np.float64)
        lib().oracle_network(ctypes.byref(self.s), _p(x), n, mode, _p(logits), _p(outs))
        return outs, logits

    def neural_query(self, nodes, rays, any_hit=False, mode=FP32, lod=0, max_steps=1000, threshold=0.5,
                     log_segments=False, segment_seed=0):
        self.s.segment_seed = segment_seed
        nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        if not log_segments:
            lib().oracle_neural_query(ctypes.byref(self.s), _p(nodes), _p(rays), len(rays), int(any_hit), mode, lod,
                                      max_steps, threshold, _p(hits), None, 0, None)
            return hits
        n_seg = ctypes.c_uint64()
        lib().oracle_neural_query(ctypes.byref(self.s), _p(nodes), _p(rays), len(rays), int(any_hit), mode, lod,
                                  max_steps, threshold, _p(hits), None, 0, ctypes.byref(n_seg))
        log = np.zeros(n_seg.value, SEG_DTYPE)
        lib().oracle_neural_query(ctypes.byref(self.s), _p(nodes), _p(rays), len(rays), int(any_hit), mode, lod,
                                  max_steps, threshold, _p(hits), _p(log), len(log), ctypes.byref(n_seg))
        return hits, log


def set_num_threads(n):
    """OpenMP threads of the oracle's parallel loops (the CPU baseline's 1-thread figure)."""
    lib().oracle_set_num_threads(int(n))


def raygen(cam12, w, h, sample=0, frame=0):
    rays = np.zeros(w * h, RAY_DTYPE)
    c = np.ascontiguousarray(cam12, np.float32)
    lib().oracle_raygen(_p(c), w, h, sample, frame, _p(rays))
    return rays


def scatter_diffuse(rays, hits, sample=0, bounce=0, verts=None, tris=None):
    """Next rays of one diffuse bounce (the restatement of the bounce kernel)."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    out = np.zeros(len(rays), RAY_DTYPE)
    v = None if verts is None else np.ascontiguousarray(verts, np.float32)
    t = None if tris is None else np.ascontiguousarray(tris, np.uint32)
    lib().oracle_scatter_diffuse(_p(v), _p(t), _p(rays), _p(hits), len(rays), sample, bounce, _p(out))
    return out


def leaf_sequence(nodes, ray, lod=0, current_maxt=np.float32(3.4028234663852886e38), cap=4096):
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    r = np.ascontiguousarray(np.asarray([ray], RAY_DTYPE))
    on = np.zeros(cap, np.uint32)
    omn = np.zeros(cap, np.float32)
    omx = np.zeros(cap, np.float32)
    n = lib().oracle_nbvh_leaf_sequence(_p(nodes), lod, _p(r), current_maxt, _p(on), _p(omn), _p(omx), cap)
    if n < 0:
        return None
    n = min(n, cap)
    return on[:n].copy(), omn[:n].copy(), omx[:n].copy()


def slab(bmin, bmax, ray):
    a, b = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
    r = np.ascontiguousarray(np.asarray([ray], RAY_DTYPE))
    tn, tf = ctypes.c_float(), ctypes.c_float()
    hit = lib().oracle_slab(_p(a), _p(b), _p(r), ctypes.byref(tn), ctypes.byref(tf))
    return bool(hit), tn.value, tf.value


def dilate(node, dilation=1e-5):
    nd = np.ascontiguousarray(np.asarray([node], NODE_DTYPE))
    a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().oracle_dilate(_p(nd), dilation, _p(a), _p(b))
    return a, b


def tri_query(verts, tris, nodes, prims, rays, any_hit=False):
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    prims = np.ascontiguousarray(prims, np.uint32)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    lib().oracle_tri_query(_p(verts), _p(tris), _p(nodes), _p(prims), _p(rays), len(rays), int(any_hit), _p(hits))
    return hits


def segment_labels(verts, tris, nodes, prims, rays, mint, maxt):
    """Network targets of segments from the triangles (generate_single_bvh_network_output_reference,
    include/cuda/neural/neural_bvh_module_common.cuh:783-861; LocalT depth :586-629): the closest
    triangle hit of each ray clipped to [mint, maxt) -> [vis, depth, normal (n+1)/2, albedo 0.9],
    zeros without a hit. Depth divides in double (the reference's 1e-3 is a double literal)."""
    r = np.array(rays, RAY_DTYPE, copy=True)
    mint = np.asarray(mint, np.float32)
    maxt = np.asarray(maxt, np.float32)
    r["tmin"], r["tmax"] = mint, maxt
    h = tri_query(verts, tris, nodes, prims, r)
    hit = (h["info"] & 0xFF) == 2
    out = np.zeros((len(r), 8), np.float32)
    out[hit, 0] = 1.0
    seg = (maxt - mint).astype(np.float32)
    num = (h["t"] - mint).astype(np.float32)
    out[hit, 1] = (num[hit].astype(np.float64) / (seg[hit].astype(np.float64) + 1e-3)).astype(np.float32)
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    t = np.asarray(tris, np.uint32).reshape(-1, 3)[h["prim_id"][hit]]
    e1, e2 = v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]
    n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1).astype(np.float32)
    inv = np.float32(1) / np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    out[hit, 2:5] = (n * inv[:, None] + np.float32(1)) * np.float32(0.5)
    out[hit, 5:8] = np.float32(0.9)
    return out


def hybrid_query(net, nbvh_nodes, verts, tris, nodes, prims, rays, any_hit=False, mode=FP32):
    """Closest (any) hit over a neural object and a triangle scene, composed from the two
    restatements like the reference's hybrid traversal resolves a ray: the triangle hit bounds the
    neural search (tmax), and a neural hit found in front of it wins."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    th = tri_query(verts, tris, nodes, prims, rays, any_hit=any_hit)
    clipped = rays.copy()
    hit = (th["info"] & 0xFF) != 0
    closer = hit & (th["t"] < clipped["tmax"])
    clipped["tmax"][closer] = th["t"][closer]
    nh = net.neural_query(nbvh_nodes, clipped, any_hit=any_hit, mode=mode)
    take = ((nh["info"] & 0xFF) != 0) & ~(any_hit & hit)
    out = th.copy()
    out[take] = nh[take]
    return out


PATH_STATE_DTYPE = np.dtype([("pixel", "<u4"), ("thr", "<f4", 3), ("illum", "<f4", 3), ("pad", "<u4")])


def path_shade(depth, rays, hits, state, accum, max_depth=5, rr_start_depth=0, sample=0, background=1,
               background_color=(0.0, 0.0, 0.0), env_albedo=(0.8, 0.8, 0.8), normal_offset=1e-4, verts=None,
               tris=None):
    """One depth of the path-trace restatement (oracle_path_shade): ending paths are written to
    accum [n_pixels, 4] in place; returns the live paths' next rays and states."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    state = np.ascontiguousarray(state, PATH_STATE_DTYPE)
    nr, ns = np.zeros(len(rays), RAY_DTYPE), np.zeros(len(rays), PATH_STATE_DTYPE)
    bg = np.asarray(background_color, np.float32)
    alb = np.asarray(env_albedo, np.float32)
    v = None if verts is None else np.ascontiguousarray(verts, np.float32)
    t = None if tris is None else np.ascontiguousarray(tris, np.uint32)
    assert accum.dtype == np.float32 and accum.flags.c_contiguous
    m = lib().oracle_path_shade(_p(v), _p(t), max_depth, rr_start_depth, sample, background, _p(bg), _p(alb),
                                normal_offset, depth, _p(rays), _p(hits), _p(state), len(rays), _p(nr), _p(ns),
                                _p(accum))
    return nr[:m].copy(), ns[:m].copy()


def path_trace(query, cam12, w, h, max_depth=5, rr_start_depth=0, sample=0, background=1,
               background_color=(0.0, 0.0, 0.0), env_albedo=(0.8, 0.8, 0.8), normal_offset=1e-4, verts=None,
               tris=None, accum=None, query_takes_depth=False):
    """One frame of 1-spp diffuse path tracing (NeuralPathModule::neural_path_tracing,
    src/cuda/neural/path/path_module.cu:139-245, diffuse materials): raygen (sample, frame 0), then
    per depth query(rays) -> hits (the oracle's neural or hybrid closest hit; query(rays, depth)
    with query_takes_depth, e.g. for a LoD switch on bounces) and path_shade.
    Returns (accum [w*h, 4], rays traced per depth)."""
    if accum is None:
        accum = np.zeros((w * h, 4), np.float32)
    rays = raygen(cam12, w, h, sample, 0)
    state = np.zeros(w * h, PATH_STATE_DTYPE)
    state["pixel"] = np.arange(w * h, dtype=np.uint32)
    state["thr"] = 1.0
    counts = []
    for depth in range(1, max_depth + 1):
        if len(rays) == 0:
            break
        counts.append(len(rays))
        hits = query(rays, depth) if query_takes_depth else query(rays)
        rays, state = path_shade(depth, rays, hits, state, accum, max_depth, rr_start_depth, sample, background,
                                 background_color, env_albedo, normal_offset, verts, tris)
    return accum, counts


class SampleCfg(ctypes.Structure):
    _fields_ = [("batch_size", ctypes.c_uint32), ("n_sub_batch", ctypes.c_uint32), ("step", ctypes.c_uint32),
                ("current_lod", ctypes.c_uint32), ("survival", ctypes.c_uint32), ("scaling", ctypes.c_float),
                ("fixed_dilation", ctypes.c_float), ("extent_dilation", ctypes.c_float),
                ("mint_eps", ctypes.c_float), ("maxt_eps", ctypes.c_float), ("survival_scale", ctypes.c_float),
                ("min_survival", ctypes.c_float), ("current_max_loss", ctypes.c_float),
                ("box_min", ctypes.c_float * 3), ("box_max", ctypes.c_float * 3), ("n_points", ctypes.c_uint32),
                ("distribution", ctypes.c_uint32)]


TREE_CUT_DTYPE = np.dtype([("loss", "<f4"), ("iter_alive", "<u4"), ("depth", "<u4"), ("samples", "<u4")])


def learn_sample(nbvh_nodes, tree_cut, verts, tris, bvh_nodes, prims, box_min, box_max, step, batch_size,
                 prev=None, n_sub_batch=1, current_lod=6, survival=True, current_max_loss=0.0, scaling=1.5,
                 fixed_dilation=1e-5, extent_dilation=1e-4, mint_eps=1e-3, maxt_eps=1e-3, survival_scale=1.0,
                 min_survival=0.005, n_points=3, distribution=1):
    """Restatement of the reference's training-data sampler (granular_sample_bvh_network_input_data,
    src/cuda/neural/neural_bvh_module_kernels.cu:119-264). prev: (inputs, targets, node_idx) the
    slots held before the step (zeros when None). Returns dict(inputs, targets, node_idx, rays,
    seg [mint, maxt], lod, dil, written) per slot."""
    c = SampleCfg(batch_size, n_sub_batch, step, current_lod, int(survival), scaling, fixed_dilation,
                  extent_dilation, mint_eps, maxt_eps, survival_scale, min_survival, current_max_loss,
                  (ctypes.c_float * 3)(*np.asarray(box_min, np.float32)),
                  (ctypes.c_float * 3)(*np.asarray(box_max, np.float32)), n_points, distribution)
    n_in = 3 * n_points
    if prev is None:
        prev = (np.zeros((batch_size, n_in), np.float32), np.zeros((batch_size, 8), np.float32),
                np.zeros(batch_size, np.uint32))
    inputs = np.array(prev[0], np.float32, copy=True).reshape(batch_size, n_in)
    targets = np.array(prev[1], np.float32, copy=True).reshape(batch_size, 8)
    node_idx = np.array(prev[2], np.uint32, copy=True)
    rays = np.zeros(batch_size, RAY_DTYPE)
    seg = np.zeros((batch_size, 2), np.float32)
    lod = np.zeros(batch_size, np.uint32)
    dil = np.zeros(batch_size, np.float32)
    written = np.zeros(batch_size, np.uint32)
    tc = np.ascontiguousarray(tree_cut, TREE_CUT_DTYPE)
    lib().oracle_learn_sample(ctypes.byref(c), _p(np.ascontiguousarray(nbvh_nodes, NODE_DTYPE)), _p(tc),
                              _p(np.ascontiguousarray(verts, np.float32)), _p(np.ascontiguousarray(tris, np.uint32)),
                              _p(np.ascontiguousarray(bvh_nodes, NODE_DTYPE)), _p(np.ascontiguousarray(prims, np.uint32)),
                              _p(inputs), _p(targets), _p(node_idx), _p(rays), _p(seg), _p(lod), _p(dil), _p(written))
    return dict(inputs=inputs, targets=targets, node_idx=node_idx, rays=rays, seg=seg, lod=lod, dil=dil,
                written=written)


def sincosf(x):
    """The sampler's float sin and cos (Cody-Waite reduction + Taylor polynomials)."""
    x = np.ascontiguousarray(x, np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib().oracle_sincosf(_p(x), len(x), _p(s), _p(c))
    return s, c


def tri_brute(verts, tris, rays):
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    lib().oracle_tri_brute(_p(verts), _p(tris), len(tris), _p(rays), len(rays), _p(hits))
    return hits


def murmur_mix(h, k):
    return lib().oracle_murmur_mix(h, k)


def murmur_fmix(h):
    return lib().oracle_murmur_fmix(h)


def lcg_seed(index, frame, linear):
    return lib().oracle_lcg_seed(index, frame, linear)


def lcg_floats(state, n):
    s = ctypes.c_uint32(state)
    return np.array([lib().oracle_lcg_float(ctypes.byref(s)) for _ in range(n)], np.float32)


def expf(x):
    return np.array([lib().oracle_expf(float(v)) for v in np.ravel(x)], np.float32).reshape(np.shape(x))


def num_threads():
    return lib().oracle_num_threads()
