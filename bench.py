#!/usr/bin/env python3
"""Benchmark: Mrays/s of the N-BVH neural ray query on one MI355X.

Workload (one step = one query over the whole ray batch, rays resident in HBM):
  displaced icosphere (6 subdivisions = 81,920 triangles, 3-octave value noise, amplitude 0.1),
  N-BVH of 256 leaves cut from its binned-SAH BVH, 1024 x 1024 jittered primary rays from a
  pinhole camera at distance 3 (fovy 45), hash grid 16 levels x 2 features, 2^19 entries per
  level, base resolution 16, scale 1.5, tables U(+-1e-4), MLP 96 -> 64 x 4 -> 8 (ReLU, Sigmoid),
  Kaiming-uniform seeded weights.
The headline value is the FP16-MFMA path; the FP32 plain path is timed beside it, the FP16
visibility decisions are compared with the FP64 oracle on a sample, and the oracle's FP32 CPU
path is timed on the host cores (cpu_baseline).

  python bench.py [--steps K] [--warmup W] [--precision fp16|bf16|fp32] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s neural query, FP16 MFMA vs FP32 plain @<=0.5% vis flip; p99 ms"
HBM_PEAK_GBS = 8000.0
MFMA_FP16_PEAK_TFLOPS = 2500.0
FLOP_PER_SEGMENT = 2 * (96 * 64 + 3 * 64 * 64 + 64 * 8)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(precision, width, height):
    """HBM bytes per launch of the query kernel on this workload, from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary (counter readings scaled by the same counters' readings for a
    device copy of known size); None when no summary covers this precision and frame size."""
    path = os.path.join(ROOT, "profiles", f"neural_query_{precision}_{width}x{height}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


_EXTRA_SCENES = {}


def sphere_field_scene(nb, ctx):
    """64 displaced geodesic spheres (2.17M triangles, seeded placement in the unit cube),
    N-BVH 4096 leaves, FP16-ready network of the headline shape (seed 2). Built once."""
    if "field" not in _EXTRA_SCENES:
        t0 = time.perf_counter()
        sc = nb.Scene(ctx, *nb.sphere_field(64, seed=5))
        sc.build_nbvh(4096)
        sc.init_network(nb.network_config(), seed=2, table_range=1e-4)
        _EXTRA_SCENES["field"] = sc
        log(f"sphere-field scene built in {time.perf_counter() - t0:.1f} s")
    return _EXTRA_SCENES["field"]


def wavefront_rays(nb, verts, tris, n_closest, n_shadow, seed=21):
    """n_closest incoherent rays (origins uniform in the unit cube, uniform directions) and
    n_shadow shadow rays from uniformly chosen triangles (uniform barycentric points, offset 1e-4
    along the face normal toward the light) to one of 8 seeded point lights, tmax = distance."""
    rng = np.random.default_rng(seed)
    rc = np.zeros(n_closest, nb.RAY_DTYPE)
    d = rng.normal(size=(n_closest, 3))
    rc["o"] = rng.uniform(0, 1, (n_closest, 3)).astype(np.float32)
    rc["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rc["tmin"], rc["tmax"] = 1e-4, np.float32(3.4028234663852886e38)
    lights = rng.uniform(-0.25, 1.25, (8, 3))
    tri = tris[rng.integers(0, len(tris), n_shadow)]
    a, b, c = verts[tri[:, 0]].astype(np.float64), verts[tri[:, 1]].astype(np.float64), verts[tri[:, 2]].astype(np.float64)
    u, v = rng.uniform(0, 1, n_shadow), rng.uniform(0, 1, n_shadow)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    p = a + (b - a) * u[:, None] + (c - a) * v[:, None]
    nrm = np.cross(b - a, c - a)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    L = lights[rng.integers(0, 8, n_shadow)]
    side = np.sign(np.einsum("ij,ij->i", L - p, nrm))[:, None]
    o = p + 1e-4 * side * nrm
    to = L - o
    dist = np.linalg.norm(to, axis=1)
    rs = np.zeros(n_shadow, nb.RAY_DTYPE)
    rs["o"] = o.astype(np.float32)
    rs["d"] = (to / dist[:, None]).astype(np.float32)
    rs["tmin"], rs["tmax"] = 1e-4, dist.astype(np.float32)
    return rc, rs


def wavefront_bench(nb, hip, ctx, steps):
    """Batched wavefront: the sphere-field scene, 4M rays in ONE persistent launch of two parts
    (2M incoherent closest-hit, 2M shadow any-hit to 8 lights), FP16 and FP32, eager and replayed
    from a captured HIP graph; the triangle-BVH fallback on the same rays beside it."""
    sc = sphere_field_scene(nb, ctx)
    cfg_v, cfg_t = nb.sphere_field(64, seed=5)
    n = 1 << 21
    rc, rs = wavefront_rays(nb, cfg_v, cfg_t, n, n)
    dc, ds = hip.Buffer.from_array(rc), hip.Buffer.from_array(rs)
    hc, hs = hip.Buffer(n * 48), hip.Buffer(n * 48)
    mod = nb.NeuralBVHModule(ctx, sc)
    parts = [(dc.ptr, n, hc.ptr, False), (ds.ptr, n, hs.ptr, True)]
    s = hip.Stream()

    def time_it(fn, k):
        fn()
        s.synchronize()
        ev = [hip.Event() for _ in range(k + 1)]
        ev[0].record(s)
        for i in range(k):
            fn()
            ev[i + 1].record(s)
        s.synchronize()
        return float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(k)]))

    out = {"workload": "64 displaced geodesic spheres, 2.17M triangles, N-BVH 4096 leaves; 4,194,304 rays in one "
                       "launch: 2M incoherent closest-hit + 2M shadow any-hit to 8 point lights; launch shape "
                       "and ray order from the autotuner on that launch",
           "rays": 2 * n}
    for name, prec in (("fp16", nb.PREC_FP16), ("fp32", nb.PREC_FP32)):
        p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec)
        # Launch shape and ray order chosen by the autotuner on the two-part launch itself (kept
        # for the scene and precision, so the launches below use them).
        shape, tune_ms = mod.autotune_parts(p, parts, stream=s.handle)
        out[name + "_autotuned_shape"] = {"blocks_per_cu": shape[0], "queue_bands": shape[1], "ray_order": shape[2],
                                          "tuning_batch_ms": round(tune_ms, 4)}
        # Default parameters: the batch has no image order, so it is traced in the order the
        # autotuner kept for the scene (the coherence order where that was faster).
        ms = time_it(lambda: mod.query_parts_device(p, parts, stream=s.handle), steps if name == "fp16" else 3)
        out[name] = {"ms": round(ms, 4), "mrays_per_s": round(2 * n / (ms * 1e-3) / 1e6, 1)}
        pa = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec, array_order=True)
        ms = time_it(lambda: mod.query_parts_device(pa, parts, stream=s.handle), steps if name == "fp16" else 3)
        out[name + "_array_order"] = {"ms": round(ms, 4), "mrays_per_s": round(2 * n / (ms * 1e-3) / 1e6, 1)}
        pw = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec, wavefront=True)
        ms = time_it(lambda: mod.query_parts_device(pw, parts, stream=s.handle), 3)
        out[name + "_wavefront_rounds"] = {"ms": round(ms, 4), "mrays_per_s": round(2 * n / (ms * 1e-3) / 1e6, 1)}
        if name == "fp16":
            h1, h2 = hc.download(nb.HIT_DTYPE), hs.download(nb.HIT_DTYPE)
            out["segments_per_ray"] = round(float((nb.hit_segments(h1).astype(np.int64).sum()
                                                   + nb.hit_segments(h2).astype(np.int64).sum()) / (2 * n)), 3)
            out["closest_hit_fraction"] = round(float((nb.hit_kind(h1) != 0).mean()), 4)
            out["shadow_occluded_fraction"] = round(float((nb.hit_kind(h2) != 0).mean()), 4)
            with hip.Graph.capture(s) as g:
                mod.query_parts_device(p, parts, stream=s.handle)
            gms = time_it(lambda: g.launch(s), steps)
            g.destroy()
            out["fp16_graph"] = {"ms": round(gms, 4), "mrays_per_s": round(2 * n / (gms * 1e-3) / 1e6, 1)}
    pt = mod.params(kind=nb.QUERY_TRIANGLE_CLOSEST)
    ms = time_it(lambda: mod.query_parts_device(pt, parts, stream=s.handle), 3)
    out["triangle_fallback"] = {"ms": round(ms, 4), "mrays_per_s": round(2 * n / (ms * 1e-3) / 1e6, 1)}
    s.destroy()
    return out


def chess_bench(nb, hip, ctx, oracle, W=1024, H=1024, steps=20):
    """The reference's own N-BVH: the shipped chess .nbvh (25,249 nodes, 12,625 leaves; a test
    fixture of this repository) with a seeded network of the shipped chess configuration (8 levels
    x 4 features, 2^18 entries, base 8, scale 2). 1024x1024 primary rays at the board; FP16, BF16
    and FP32 query time, and visibility flips of the FP16 network vs FP64 on every 16th row."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                        "chess_nodes_32940_hash_18_lvl_8.nbvh")
    cfg = nb.network_config(n_levels=8, n_features_per_level=4, log2_hashmap_size=18, base_resolution=8,
                            per_level_scale=2.0)
    sc = nb.Scene(ctx)
    sc.load_neural_bvh(path)
    sc.init_network(cfg, seed=9, table_range=1e-2)
    lo, hi = sc.bbox()
    c = (lo + hi) * 0.5
    ext = float(np.max(hi - lo))
    cam = nb.camera_look_at(tuple(c + np.array([0.0, 0.9, 1.1]) * ext), tuple(c), (0, 1, 0), 45.0, W, H)
    d_rays, d_hits = hip.Buffer(W * H * 32), hip.Buffer(W * H * 48)
    ctx.raygen(cam, W, H, d_rays.ptr)
    mod = nb.NeuralBVHModule(ctx, sc)
    s = hip.Stream()
    out = {"workload": f"reference chess N-BVH (25,249 nodes), hash grid 8x4 2^18 base 8 scale 2, MLP 96-64x4-8, "
                       f"{W}x{H} primary rays"}
    for name, prec, k in (("fp16", nb.PREC_FP16, steps), ("bf16", nb.PREC_BF16, steps), ("fp32", nb.PREC_FP32, 5)):
        p = mod.params(precision=prec, image_width=W)
        mod.query_device(p, d_rays.ptr, W * H, d_hits.ptr, stream=s.handle)
        ev = [hip.Event() for _ in range(k + 1)]
        ev[0].record(s)
        for i in range(k):
            mod.query_device(p, d_rays.ptr, W * H, d_hits.ptr, stream=s.handle)
            ev[i + 1].record(s)
        s.synchronize()
        ms = float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(k)]))
        hits = d_hits.download(nb.HIT_DTYPE)
        out[name] = {"ms": round(ms, 4), "mrays_per_s": round(W * H / (ms * 1e-3) / 1e6, 1),
                     "segments_per_ray": round(float(nb.hit_segments(hits).mean()), 3),
                     "hit_fraction": round(float((nb.hit_kind(hits) != 0).mean()), 4)}
    out["fp16_speedup_vs_fp32"] = round(out["fp32"]["ms"] / out["fp16"]["ms"], 3)
    cfg2, tables, mlp = sc.network()
    net = oracle.Net(cfg2, tables, mlp, lo, hi)
    sample = d_rays.download(nb.RAY_DTYPE).reshape(H, W)[::16].reshape(-1)
    _, seglog = net.neural_query(sc.nbvh(), sample, mode=oracle.FP64, log_segments=True)
    x = net.encode(sample[seglog["ray"]], seglog["mint"], seglog["maxt"], mode=oracle.FP32)
    _, z = mod.network_inference(x, precision=nb.PREC_FP16)
    out["fp16_vis_flip_rate"] = round(float(np.mean((z[:, 0] > 0) != (seglog["logit"][:, 0] > 0))), 5)
    out["flip_sample_segments"] = int(len(x))
    s.destroy()
    sc.close()
    log(f"chess N-BVH: fp16 {out['fp16']['mrays_per_s']} / bf16 {out['bf16']['mrays_per_s']} / fp32 "
        f"{out['fp32']['mrays_per_s']} Mrays/s, speedup {out['fp16_speedup_vs_fp32']}x, "
        f"flips {out['fp16_vis_flip_rate']}")
    return out


def train_bench(nb, hip, ctx, oracle, batch=1 << 16, steps=100, pool=1 << 20):
    """Training step throughput (nbvh_trainer_step, FP32) of the headline network on the
    benchmark mesh: segments sampled in its N-BVH leaves (nbvh_amd.leaf_segments), encoded and
    labelled from the triangles, minibatches of `batch` per step; the held-out visibility accuracy
    before and after."""
    v, t = nb.icosphere(6, seed=7, amplitude=0.1)
    sc = nb.Scene(ctx, v, t)
    sc.build_nbvh(256)
    sc.init_network(nb.network_config(), seed=1, table_range=1e-2)
    cfg, tables, mlp = sc.network()
    lo, hi = sc.bbox()
    net = oracle.Net(cfg, tables, mlp, lo, hi)
    rays, mint, maxt = nb.leaf_segments(sc.nbvh(), pool + 20000, seed=3)
    x = net.encode(rays, mint, maxt, mode=oracle.FP32)
    mod = nb.NeuralBVHModule(ctx, sc)
    tg = mod.segment_labels(rays, mint, maxt)
    held = slice(0, 20000)

    def accuracy():
        out, _ = mod.network_inference(x[held], precision=nb.PREC_FP32)
        return float(np.mean((out[:, 0] > 0.5) == (tg[held, 0] > 0.5)))

    acc0 = accuracy()
    dx, dt = hip.Buffer.from_array(x[20000:]), hip.Buffer.from_array(tg[20000:])
    tr = nb.Trainer(ctx, sc, max_batch=batch)
    s = hip.Stream()
    rng = np.random.default_rng(5)
    offs = rng.integers(0, pool - batch, steps + 1)
    first = tr.step_device(dx.ptr + int(offs[0]) * 36, dt.ptr + int(offs[0]) * 32, batch, stream=s.handle)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for k in range(1, steps + 1):  # queued back to back: no loss read, no host round trip per step
        tr.step_device(dx.ptr + int(offs[k]) * 36, dt.ptr + int(offs[k]) * 32, batch, stream=s.handle,
                       want_loss=False)
    e1.record(s)
    s.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    last = tr.step_device(dx.ptr, dt.ptr, batch, gradients_only=True, stream=s.handle)
    tr.commit()
    acc1 = accuracy()
    tr.close()
    s.destroy()
    sc.close()
    # The table gradient has one term per (point, level, corner, feature) of every sample. Added
    # with global FP32 atomics, the GPU's random atomic-add rate would bound the step from below;
    # the trainer instead sums each slice of the gradient in LDS (no global atomics but for the
    # split slices of the coarse levels), and this floor shows what that avoids.
    adds = batch * cfg.n_points * cfg.grid.n_levels * 8 * cfg.grid.n_features_per_level
    ceiling = atomic_ceiling()
    out = {"workload": f"headline network (hash grid 16x2 2^19, MLP 96-64x4-8) on the icosphere's 256 N-BVH leaves, "
                       f"FP32, minibatches of {batch} leaf segments from a pool of {pool}",
           "ms_per_step": round(ms, 3), "samples_per_s": round(batch / (ms * 1e-3) / 1e6, 2),
           "unit": "M samples/s", "steps": steps, "loss_first": round(first, 5), "loss_last": round(last, 5),
           "heldout_visibility_accuracy": {"before": round(acc0, 4), "after": round(acc1, 4),
                                           "class_prior": round(float(max(tg[held, 0].mean(),
                                                                              1 - tg[held, 0].mean())), 4)},
           "table_gradient": {"terms_per_step": adds, "gpu_atomic_adds_per_s": round(ceiling, 1),
                              "global_atomics_floor_ms": round(adds / ceiling * 1e3, 3),
                              "step_vs_that_floor": round(ms / (adds / ceiling * 1e3), 3)}}
    # Step time with larger hash tables (same segments): from 2^21 entries per level the hashed
    # levels' table gradient is binned, below that scanned once per slice.
    big = {}
    for log2 in (20, 21, 22, 23):
        sc2 = nb.Scene(ctx, v, t)
        sc2.build_nbvh(256)
        sc2.init_network(nb.network_config(log2_hashmap_size=log2), seed=1, table_range=1e-2)
        tr2 = nb.Trainer(ctx, sc2, max_batch=batch)
        s2 = hip.Stream()
        tr2.step_device(dx.ptr, dt.ptr, batch, stream=s2.handle, want_loss=False)
        f0, f1 = hip.Event(), hip.Event()
        f0.record(s2)
        for k in range(1, 11):
            tr2.step_device(dx.ptr + int(offs[k]) * 36, dt.ptr + int(offs[k]) * 32, batch, stream=s2.handle,
                            want_loss=False)
        f1.record(s2)
        s2.synchronize()
        big[f"2^{log2}"] = round(f0.elapsed_ms(f1) / 10, 3)
        tr2.close()
        s2.destroy()
        sc2.close()
    out["ms_per_step_by_table_size"] = dict({"2^19": out["ms_per_step"]}, **big)
    log(f"training: {out['ms_per_step']} ms per step of {batch} ({out['samples_per_s']} M samples/s), loss "
        f"{out['loss_first']} -> {out['loss_last']}, held-out visibility accuracy {acc0:.3f} -> {acc1:.3f}; "
        f"the {adds / 1e6:.1f} M table-gradient terms as global atomic adds ({ceiling / 1e9:.1f} G adds/s) "
        f"would take {adds / ceiling * 1e3:.2f} ms alone")
    return out


def mlp_mfma_bench(nb, hip, ctx, n=1 << 21, steps=10):
    """The network kernel (nbvh_network_inference: hash-grid gathers + the 96-64x4-8 MLP) on n
    encoded segments with tables of 2^14 entries per level (cache-resident, 16 levels x 2
    features), so that the MLP rather than the gathers sets the time: MLP TFLOP/s (37,888 FLOP per
    segment) against MI355X's dense peaks (FP16/BF16 2.5 PFLOP/s on MFMA, FP32 157.3 TFLOP/s)."""
    v, t = nb.icosphere(3, seed=7, amplitude=0.1)
    sc = nb.Scene(ctx, v, t)
    sc.build_nbvh(32)
    sc.init_network(nb.network_config(log2_hashmap_size=14), seed=1, table_range=1e-2)
    mod = nb.NeuralBVHModule(ctx, sc)
    x = np.random.default_rng(3).uniform(0.05, 0.95, (n, 9)).astype(np.float32)
    d_in, d_out = hip.Buffer.from_array(x), hip.Buffer(n * 8 * 4)
    s = hip.Stream()
    out = {"workload": f"{n} segments, hash grid 16 x 2 at 2^14 entries per level (cache-resident), MLP 96-64x4-8"}
    for name, prec, peak in (("fp16", nb.PREC_FP16, 2500.0), ("bf16", nb.PREC_BF16, 2500.0), ("fp32", nb.PREC_FP32, 157.3)):
        mod.network_inference_device(prec, d_in.ptr, n, d_out.ptr, stream=s.handle)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(s)
        for _ in range(steps):
            mod.network_inference_device(prec, d_in.ptr, n, d_out.ptr, stream=s.handle)
        e1.record(s)
        s.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        tf = n * FLOP_PER_SEGMENT / (ms * 1e-3) / 1e12
        out[name] = {"ms": round(ms, 4), "mlp_tflops": round(tf, 1), "frac_of_dense_peak": round(tf / peak, 4)}
    s.destroy()
    sc.close()
    log("mlp on MFMA (cache-resident tables): " + ", ".join(
        f"{k} {out[k]['mlp_tflops']} TFLOP/s ({out[k]['frac_of_dense_peak']:.1%} of peak)" for k in ("fp16", "bf16", "fp32")))
    return out


def atomic_ceiling(entries=6529022, waves_per_cu=8, per_lane=192):
    """Rate of random FP32 atomic adds (pairs on the two floats of one entry, as the table-gradient
    scatter issues them) into a table of the headline grid's size (nbvh_amd/libnbvh_bench.so,
    csrc/bench/gather_ceiling.hip); the table's final sum is checked against the number of adds."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "nbvh_amd", "libnbvh_bench.so"))
    fn = lib.nbvh_bench_atomic_adds
    fn.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
    fn.restype = ctypes.c_int
    ms, rate, exact = ctypes.c_float(), ctypes.c_double(), ctypes.c_int32()
    st = fn(entries, waves_per_cu, per_lane, ctypes.byref(ms), ctypes.byref(rate), ctypes.byref(exact))
    if st != 0 or exact.value != 1:
        raise RuntimeError(f"atomic-add ceiling failed (HIP error {st}, exact sum {exact.value})")
    return rate.value


def gather_ceiling(sizes_mib=(2, 4, 8, 16, 32, 64)):
    """Random-gather line rate of the GPU (nbvh_amd/libnbvh_bench.so, csrc/bench/gather_ceiling.hip):
    8 waves per CU, every lane of a load instruction in its own random 128-byte line of a table of
    the given size. The ceiling the query's fine-level hash-grid gathers run against."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "nbvh_amd", "libnbvh_bench.so"))
    fn = lib.nbvh_bench_gather_lines
    fn.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    out = {"workload": "8 waves per CU, 64 distinct random 128-B lines per load instruction, 512 loads per lane"}
    for mib in sizes_mib:
        ms, rate = ctypes.c_float(), ctypes.c_double()
        st = fn(mib << 20, 1, 8, 512, ctypes.byref(ms), ctypes.byref(rate))
        if st != 0:
            raise RuntimeError(f"gather ceiling failed with HIP error {st}")
        out[f"{mib}MiB_glines_per_s"] = round(rate.value / 1e9, 1)
    log("random-gather line rate (G lines/s): " + ", ".join(f"{k.split('MiB')[0]} MiB {v}" for k, v in out.items()
                                                           if k.endswith("per_s")))
    return out


def plain_triangles(nb, hip, ctx, scene, oracle, W=512, H=512):
    """The plain non-neural workload: 512x512 jittered primary rays, closest hit on the triangle BVH
    (binned SAH, leaf <= 2) of the 81,920-triangle icosphere. GPU kernel vs the oracle's CPU
    traversal (OpenMP, all host threads); triangle ids and t must be bit-identical."""
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    rays = oracle.raygen(nb.camera_array(cam), W, H)
    v, t = nb.icosphere(6, seed=7, amplitude=0.1)
    nodes, prims = scene.bvh()
    mod = nb.NeuralBVHModule(ctx, scene)
    dr, dh = hip.Buffer.from_array(rays), hip.Buffer(len(rays) * 48)
    p = mod.params(kind=nb.QUERY_TRIANGLE_CLOSEST)
    s = hip.Stream()
    mod.query_device(p, dr.ptr, len(rays), dh.ptr, stream=s.handle)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(s)
    for _ in range(10):
        mod.query_device(p, dr.ptr, len(rays), dh.ptr, stream=s.handle)
    e1.record(s)
    s.synchronize()
    gpu_ms = e0.elapsed_ms(e1) / 10
    got = dh.download(nb.HIT_DTYPE)
    s.destroy()
    oracle.tri_query(v, t, nodes, prims, rays[:1024])
    reps, t0 = 0, time.perf_counter()
    while True:
        exp = oracle.tri_query(v, t, nodes, prims, rays)
        reps += 1
        el = time.perf_counter() - t0
        if reps >= 5 or el > 10.0:
            break
    same = bool(np.array_equal(got["prim_id"], exp["prim_id"]) and
                np.array_equal(got["t"].view(np.uint32), exp["t"].view(np.uint32)))
    out = {"rays": len(rays), "gpu_mrays_per_s": round(len(rays) / (gpu_ms * 1e-3) / 1e6, 1),
           "cpu_mrays_per_s": round(reps * len(rays) / el / 1e6, 2), "cpu_threads": oracle.num_threads(),
           "prim_id_and_t_bit_exact": same, "hit_fraction": round(float((nb.hit_kind(got) != 0).mean()), 4)}
    log(f"plain triangles 512x512: GPU {out['gpu_mrays_per_s']} Mrays/s, CPU {out['cpu_mrays_per_s']} Mrays/s "
        f"({out['cpu_threads']} threads), bit-exact {same}")
    return out


def _sig(z):
    return 1.0 / (1.0 + np.exp(-np.asarray(z, np.float64)))


def sig4(v):
    """A reported error to 4 significant digits (never rounded to zero)."""
    return float(f"{float(v):.4g}")


def output_errors(z, z64):
    """Per-output error of network logits z against FP64 logits z64, in what the query uses:
    visibility decision flips; |depth - depth64| (sigmoid output, LocalT fraction of the segment);
    normal angle (degrees); |albedo - albedo64| (sigmoid outputs); and the logits' error relative
    to each output's range over the sample (p99 over segments, worst output of each group)."""
    s, s64 = _sig(z), _sig(z64)
    n = 2 * s[:, 2:5] - 1
    n64 = 2 * s64[:, 2:5] - 1
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    n64 /= np.maximum(np.linalg.norm(n64, axis=1, keepdims=True), 1e-30)
    ang = np.degrees(np.arccos(np.clip(np.einsum("ij,ij->i", n, n64), -1, 1)))
    dd = np.abs(s[:, 1] - s64[:, 1])
    da = np.abs(s[:, 5:8] - s64[:, 5:8]).max(axis=1)
    rng = np.maximum(z64.max(axis=0) - z64.min(axis=0), 1e-30)
    rel = np.percentile(np.abs(np.asarray(z, np.float64) - z64) / rng, 99, axis=0)
    out = {"vis_flip_rate": float(np.mean((z[:, 0] > 0) != (z64[:, 0] > 0))),
           "depth_abs_p99": float(np.percentile(dd, 99)), "depth_abs_max": float(dd.max()),
           "normal_deg_p99": float(np.percentile(ang, 99)), "albedo_abs_p99": float(np.percentile(da, 99)),
           "logit_rel_range_p99": {"visibility": float(rel[0]), "depth": float(rel[1]), "normal": float(rel[2:5].max()),
                                   "albedo": float(rel[5:8].max())}}
    return {k: ({kk: sig4(vv) for kk, vv in v.items()} if isinstance(v, dict) else sig4(v)) for k, v in out.items()}


def path_traced_bench(nb, hip, ctx, scene, oracle, W=1920, H=1080, bounces=4, every=64):
    """Precision workload: 1920x1080 primary rays (1 spp) and 4 diffuse bounces. Paths continue
    from the FP16 query's hits through nbvh_scatter_diffuse (cosine-sampled directions from the
    reference's LCG); rays that missed end. Per bounce: FP16 and FP32 query times over that
    bounce's rays, and on every 64th ray the FP64 oracle's segments evaluated by the FP16 / BF16 /
    FP8 networks -> per-output errors vs FP64; the FP16 query's hit/miss decisions vs the FP64
    oracle query on the same rays."""
    mod = nb.NeuralBVHModule(ctx, scene)
    cfg, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    net = oracle.Net(cfg, tables, mlp, bmin, bmax)
    nodes = scene.nbvh()
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    n = W * H
    d_rays, d_hits, d_next = hip.Buffer(n * 32), hip.Buffer(n * 48), hip.Buffer(n * 32)
    s = hip.Stream()
    ctx.raygen(cam, W, H, d_rays.ptr, stream=s.handle)
    s.synchronize()

    def timed(prec, k, cur, sort_rays=False):
        p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec, sort_rays=sort_rays, array_order=not sort_rays)
        mod.query_device(p, d_rays.ptr, cur, d_hits.ptr, stream=s.handle)
        e = [hip.Event() for _ in range(k + 1)]
        e[0].record(s)
        for i in range(k):
            mod.query_device(p, d_rays.ptr, cur, d_hits.ptr, stream=s.handle)
            e[i + 1].record(s)
        s.synchronize()
        return float(np.mean([e[i].elapsed_ms(e[i + 1]) for i in range(k)]))

    rows, worst = [], {}
    cur = n
    for b in range(bounces + 1):
        # Bounces are incoherent: they are traced in the sorted coherence order (same hits).
        srt = b > 0
        ms32 = timed(nb.PREC_FP32, 2, cur, srt)
        ms16 = timed(nb.PREC_FP16, 5, cur, srt)  # leaves the FP16 hits in d_hits
        rays = d_rays.download(nb.RAY_DTYPE, cur)
        hits = d_hits.download(nb.HIT_DTYPE, cur)
        samp = rays[::every]
        e64, seglog = net.neural_query(nodes, samp, mode=oracle.FP64, log_segments=True)
        row = {"bounce": b, "rays": cur, "sorted": srt, "fp16_mrays_per_s": round(cur / (ms16 * 1e-3) / 1e6, 1),
               "fp32_mrays_per_s": round(cur / (ms32 * 1e-3) / 1e6, 1),
               "fp16_query_kind_mismatch": round(float(np.mean(nb.hit_kind(hits[::every]) != nb.hit_kind(e64))), 5),
               "sample_segments": int(len(seglog))}
        if len(seglog):
            x = net.encode(samp[seglog["ray"]], seglog["mint"], seglog["maxt"], mode=oracle.FP32)
            z64 = seglog["logit"].astype(np.float64)
            for name, prec in (("fp16", nb.PREC_FP16), ("bf16", nb.PREC_BF16), ("fp8", nb.PREC_FP8)):
                _, z = mod.network_inference(x, precision=prec)
                err = output_errors(z, z64)
                row[name] = err
                w = worst.setdefault(name, {})
                for k, v in err.items():
                    if isinstance(v, dict):
                        wk = w.setdefault(k, {})
                        for kk, vv in v.items():
                            wk[kk] = max(wk.get(kk, 0.0), vv)
                    else:
                        w[k] = max(w.get(k, 0.0), v)
        rows.append(row)
        log_msg = f"path-traced bounce {b}: {cur} rays, fp16 {row['fp16_mrays_per_s']} / fp32 " \
                  f"{row['fp32_mrays_per_s']} Mrays/s"
        if "fp16" in row:
            log_msg += f", flips fp16 {row['fp16']['vis_flip_rate']} bf16 {row['bf16']['vis_flip_rate']} " \
                       f"fp8 {row['fp8']['vis_flip_rate']}"
        log(log_msg)
        if b == bounces:
            break
        mod.scatter_diffuse_device(d_rays.ptr, d_hits.ptr, cur, d_next.ptr, sample=0, bounce=b, stream=s.handle)
        s.synchronize()
        nxt = d_next.download(nb.RAY_DTYPE, cur)
        nxt = nxt[nxt["tmax"] > 0]
        cur = len(nxt)
        if cur == 0:
            break
        d_rays.upload(nxt)
    s.destroy()
    return {"workload": f"{W}x{H} primary rays, 1 spp, {bounces} diffuse bounces from the FP16 query's hits "
                        "(misses end the path), icosphere 81,920 tris, N-BVH 256, headline network; errors on "
                        f"every {every}th ray's segments vs the FP64 oracle",
            "per_bounce": rows, "worst_over_bounces": worst}


def hybrid_environment(nb):
    """Triangle meshes around the neural icosphere: a 6x6 floor quad at y = -1.3 and four
    displaced geodesic spheres (frequency 12, radius 0.45)."""
    vs = [np.array([[-3, -1.3, -3], [3, -1.3, -3], [3, -1.3, 3], [-3, -1.3, 3]], np.float32)]
    ts = [np.array([[0, 2, 1], [0, 3, 2]], np.uint32)]
    off = 4
    for i, c in enumerate([(1.6, -0.8, 0.2), (-1.5, -0.7, -0.4), (0.3, 1.5, -0.9), (-0.2, -0.9, 1.6)]):
        v, t = nb.geosphere(12, seed=40 + i, amplitude=0.05, center=c, radius=0.45)
        vs.append(v)
        ts.append(t + np.uint32(off))
        off += len(v)
    return np.concatenate(vs), np.concatenate(ts)


def hybrid_path_traced_bench(nb, hip, ctx, scene, W=1920, H=1080, bounces=4):
    """Hybrid paths: the neural icosphere inside a triangle environment (hybrid_environment),
    1920x1080 primary rays and 4 diffuse bounces, every bounce a hybrid closest-hit query (triangles
    first, the neural object in front of them); FP16 and FP32 timed per bounce."""
    env = nb.Scene(ctx, *hybrid_environment(nb))
    mod = nb.NeuralBVHModule(ctx, scene)
    cam = nb.camera_look_at((0.5, 0.6, 3.2), (0, -0.2, 0), (0, 1, 0), 45.0, W, H)
    n = W * H
    d_rays, d_hits, d_next = hip.Buffer(n * 32), hip.Buffer(n * 48), hip.Buffer(n * 32)
    s = hip.Stream()
    ctx.raygen(cam, W, H, d_rays.ptr, stream=s.handle)
    s.synchronize()
    env_mod = nb.NeuralBVHModule(ctx, env)

    def timed(prec, k, cur, srt):
        p = mod.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec, sort_rays=srt, array_order=not srt)
        mod.query_hybrid_device(env, p, d_rays.ptr, cur, d_hits.ptr, stream=s.handle)
        e = [hip.Event() for _ in range(k + 1)]
        e[0].record(s)
        for i in range(k):
            mod.query_hybrid_device(env, p, d_rays.ptr, cur, d_hits.ptr, stream=s.handle)
            e[i + 1].record(s)
        s.synchronize()
        return float(np.mean([e[i].elapsed_ms(e[i + 1]) for i in range(k)]))

    rows, cur = [], n
    for b in range(bounces + 1):
        srt = b > 0
        ms32 = timed(nb.PREC_FP32, 2, cur, srt)
        ms16 = timed(nb.PREC_FP16, 5, cur, srt)
        hits = d_hits.download(nb.HIT_DTYPE, cur)
        k = nb.hit_kind(hits)
        rows.append({"bounce": b, "rays": cur, "sorted": srt, "fp16_mrays_per_s": round(cur / (ms16 * 1e-3) / 1e6, 1),
                     "fp32_mrays_per_s": round(cur / (ms32 * 1e-3) / 1e6, 1),
                     "neural_hits": round(float((k == nb.HIT_NEURAL).mean()), 4),
                     "triangle_hits": round(float((k == nb.HIT_TRIANGLE).mean()), 4)})
        if b == bounces:
            break
        env_mod.scatter_diffuse_device(d_rays.ptr, d_hits.ptr, cur, d_next.ptr, sample=0, bounce=b, stream=s.handle)
        s.synchronize()
        nxt = d_next.download(nb.RAY_DTYPE, cur)
        nxt = nxt[nxt["tmax"] > 0]
        cur = len(nxt)
        if cur == 0:
            break
        d_rays.upload(nxt)
    s.destroy()
    env.close()
    log("hybrid path-traced: " + ", ".join(f"b{r['bounce']} {r['fp16_mrays_per_s']}/{r['fp32_mrays_per_s']}"
                                           for r in rows) + " Mrays/s (fp16/fp32)")
    return {"workload": f"{W}x{H} primary rays + {bounces} diffuse bounces, hybrid closest hit: neural icosphere "
                        "(N-BVH 256, headline network) inside a floor quad and 4 geodesic spheres (triangles)",
            "per_bounce": rows}


def path_trace_driver_bench(nb, hip, ctx, scene, W=1920, H=1080, max_depth=5, rr_start_depth=3, frames=3):
    """Native path-trace driver (nbvh_path_trace): 1920x1080 at 1 spp, 5 path vertices (4 diffuse
    bounces), Russian roulette from depth 3, sky background, device-side compaction, bounces in
    coherence order. Neural icosphere alone and inside the triangle environment (hybrid).
    Frame time with HIP events around the whole call (live-path counts stay on the device; the
    per-depth counts are read back once at the end), and of the same frame captured into a HIP
    graph (caller workspace, no read-back) and replayed; the replay must equal the eager frame."""
    env = nb.Scene(ctx, *hybrid_environment(nb))
    mod = nb.NeuralBVHModule(ctx, scene)
    acc = hip.Buffer(W * H * 16)
    s = hip.Stream()
    out = {"workload": f"{W}x{H}, 1 spp, max_depth {max_depth} (4 diffuse bounces), Russian roulette from depth "
                       f"{rr_start_depth}, sky background; 'neural': icosphere 81,920 tris as N-BVH 256 + headline "
                       "network; 'hybrid': the same inside a floor quad and 4 geodesic spheres (triangles)"}
    for scene_name, cam, environment in (
            ("neural", nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H), None),
            ("hybrid", nb.camera_look_at((0.5, 0.6, 3.2), (0, -0.2, 0), (0, 1, 0), 45.0, W, H), env)):
        for name, prec in (("fp16", nb.PREC_FP16), ("fp32", nb.PREC_FP32)):
            kw = dict(environment=environment, precision=prec, max_depth=max_depth, rr_start_depth=rr_start_depth,
                      stream=s.handle)
            mod.path_trace_device(cam, W, H, acc.ptr, **kw)
            s.synchronize()
            ms, counts = [], None
            for f in range(frames):
                e0, e1 = hip.Event(), hip.Event()
                e0.record(s)
                counts = mod.path_trace_device(cam, W, H, acc.ptr, sample=f, **kw)
                e1.record(s)
                s.synchronize()
                ms.append(e0.elapsed_ms(e1))
            img = acc.download(np.float32).reshape(-1, 4)
            m = float(np.median(ms))
            out[f"{scene_name}_{name}"] = {"ms_per_frame": round(m, 3), "rays_per_depth": counts,
                                           "mrays_per_s": round(sum(counts) / (m * 1e-3) / 1e6, 1),
                                           "mean_radiance": round(float((img[:, :3] / img[:, 3:4]).mean()), 5)}
            # The frame of sample 0 recorded into one HIP graph and replayed.
            pk = dict(precision=prec, max_depth=max_depth, rr_start_depth=rr_start_depth, sample=0)
            ws = hip.Buffer(mod.path_trace_workspace_size(mod.path_params(W, H, **pk), W, H, environment))
            pg = mod.path_params(W, H, workspace=ws, **pk)
            mod.path_trace_params_device(pg, cam, W, H, acc.ptr, environment, stream=s.handle)
            s.synchronize()
            ref = acc.download(np.float32)
            with hip.Graph.capture(s) as g:
                mod.path_trace_params_device(pg, cam, W, H, acc.ptr, environment, stream=s.handle, stats=False)
            ev = [hip.Event() for _ in range(frames + 1)]
            ev[0].record(s)
            for f in range(frames):
                g.launch(s)
                ev[f + 1].record(s)
            s.synchronize()
            gms = float(np.median([ev[f].elapsed_ms(ev[f + 1]) for f in range(frames)]))
            out[f"{scene_name}_{name}"]["graph_ms_per_frame"] = round(gms, 3)
            out[f"{scene_name}_{name}"]["graph_equals_eager"] = bool(np.array_equal(acc.download(np.float32), ref))
            g.destroy()
            ws.free()
    s.destroy()
    env.close()
    log("path-trace driver: " + ", ".join(f"{k} {v['ms_per_frame']} ms ({v['mrays_per_s']} Mrays/s, graph "
                                          f"{v['graph_ms_per_frame']} ms)" for k, v in out.items()
                                          if isinstance(v, dict)))
    return out


def serving_bench(nb, hip, ctx, scene_a, seconds, load, seed=11, n_streams=4, max_batch_rays=1 << 21,
                  host_fraction=0.25):
    """Continuous-batching serving run over three resident scenes.

    Scenes: the 81,920-triangle icosphere (N-BVH 256), a field of 64 geodesic spheres (2.17M
    triangles, N-BVH 4096) and one geodesic sphere of frequency 648 (8.4M triangles, N-BVH
    16384), each with its own seeded FP16 network of the headline shape. Requests arrive as a
    Poisson process; sizes are lognormal (median 65,536 rays, sigma 1) clamped to [4,096, 2M];
    70 % closest-hit, 30 % any-hit (shadow) neural queries; the scene is drawn uniformly. Rays
    are windows of a 2048x1024 camera frame of the request's scene: resident in HBM with hits to a
    device ring (zero-copy device requests), or, for `host_fraction` of the requests, in host
    memory with hits to a host ring (staged through the server's arena, admitted by its
    capacity). The arrival rate is set to `load` times the measured single-launch capacity of the
    mix. Latency is submit -> hits complete (host requests: hits copied back), from the server's
    own steady-clock timestamps."""
    import threading
    rng = np.random.default_rng(seed)
    log("serving: building the 8.4M-triangle scene")
    t0 = time.perf_counter()
    big = nb.Scene(ctx, *nb.geosphere(648, seed=9, amplitude=0.1))
    big.build_nbvh(16384)
    big.init_network(nb.network_config(), seed=3, table_range=1e-4)
    scenes = [scene_a, sphere_field_scene(nb, ctx), big]
    log(f"serving: scenes ready in {time.perf_counter() - t0:.1f} s")
    W, H = 2048, 1024
    pool = []
    for sc in scenes:
        lo, hi = sc.bbox()
        c = (lo + hi) * 0.5
        ext = float(np.max(hi - lo))
        cam = nb.camera_look_at(tuple(c + np.array([0.3, 0.4, 1.0]) * 1.6 * ext), tuple(c), (0, 1, 0), 45.0, W, H)
        d = hip.Buffer(W * H * 32)
        ctx.raygen(cam, W, H, d.ptr)
        pool.append(d)
    hip.synchronize()
    # Host-resident requests use pinned memory, as a serving front end would.
    pinned = [hip.HostArray(nb.RAY_DTYPE, W * H) for _ in pool]
    for d, h in zip(pool, pinned):
        h.array[:] = d.download(nb.RAY_DTYPE)
    host_pool = [h.array for h in pinned]
    ring_rays = 16 << 20
    d_hits = hip.Buffer(ring_rays * 48)
    host_ring = 4 << 20
    pinned_hits = hip.HostArray(nb.HIT_DTYPE, host_ring)
    h_hits = pinned_hits.array

    def params(kind):
        return nb.NeuralBVHModule.params(kind=kind, precision=nb.PREC_FP16)

    # Capacity of the mix: one 1M-ray launch per scene and kind, timed with events.
    s = hip.Stream()
    cap = []
    for sc, d in zip(scenes, pool):
        mod = nb.NeuralBVHModule(ctx, sc)
        for kind in (nb.QUERY_NEURAL_CLOSEST, nb.QUERY_NEURAL_ANY):
            p = params(kind)
            mod.query_device(p, d.ptr, 1 << 20, d_hits.ptr, stream=s.handle)
            e0, e1 = hip.Event(), hip.Event()
            e0.record(s)
            mod.query_device(p, d.ptr, 1 << 20, d_hits.ptr, stream=s.handle)
            e1.record(s)
            s.synchronize()
            cap.append((1 << 20) / (e0.elapsed_ms(e1) * 1e-3))
    s.destroy()
    # Mix weights: scene uniform, 70/30 kinds -> time per ray is the weighted mean of 1/capacity.
    sec_per_ray = np.mean([0.7 / cap[2 * i] + 0.3 / cap[2 * i + 1] for i in range(3)])
    capacity = 1.0 / sec_per_ray
    # Host requests also move 32 B in and 48 B out per ray over PCIe: measured as the time of a
    # 1M-ray host request through a server alone, minus its device-request time.
    with nb.Server(ctx, n_streams=1, max_batch_rays=max_batch_rays) as probe:
        p0 = params(nb.QUERY_NEURAL_CLOSEST)
        probe.wait(probe.submit(scenes[0], p0, host_pool[0][:1 << 20], h_hits[:1 << 20]))
        t0 = time.perf_counter()
        probe.wait(probe.submit(scenes[0], p0, host_pool[0][:1 << 20], h_hits[:1 << 20]))
        t_host = time.perf_counter() - t0
        t0 = time.perf_counter()
        probe.wait(probe.submit_device(scenes[0], p0, pool[0].ptr, 1 << 20, d_hits.ptr))
        t_dev = time.perf_counter() - t0
    copy_sec_per_ray = max(0.0, t_host - t_dev) / (1 << 20)
    capacity = 1.0 / (sec_per_ray + host_fraction * copy_sec_per_ray)

    sizes = np.clip(rng.lognormal(np.log(65536), 1.0, 200000), 4096, 2 << 20).astype(np.int64)
    mean_size = float(sizes.mean())
    rate = load * capacity / mean_size
    n_req = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n_req)
    reqs = []
    for i in range(n_req):
        n = int(sizes[i])
        sc = int(rng.integers(3))
        kind = nb.QUERY_NEURAL_CLOSEST if rng.random() < 0.7 else nb.QUERY_NEURAL_ANY
        off = int(rng.integers(0, W * H - n + 1))
        reqs.append((sc, kind, off, n, bool(rng.random() < host_fraction)))
    log(f"serving: capacity of the mix {capacity / 1e6:.0f} Mrays/s, mean request {mean_size:.0f} rays, "
        f"{rate:.0f} requests/s for {seconds:.1f} s ({n_req} requests, load {load})")

    srv = nb.Server(ctx, n_streams=n_streams, max_batch_rays=max_batch_rays)
    # Warm-up: every scene x kind once through the server.
    for i in range(3):
        for kind in (nb.QUERY_NEURAL_CLOSEST, nb.QUERY_NEURAL_ANY):
            srv.wait(srv.submit_device(scenes[i], params(kind), pool[i].ptr, 65536, d_hits.ptr))
    tickets = []
    cond = threading.Condition()
    done = [False]
    infos = []

    def collector():
        i = 0
        while True:
            with cond:
                while i >= len(tickets) and not done[0]:
                    cond.wait()
                if i >= len(tickets):
                    return
                t, n, host = tickets[i]
            _, info = srv.wait(t)
            infos.append((info, n, host))
            i += 1

    th = threading.Thread(target=collector)
    th.start()
    ring, hring = 0, 0
    ps = {k: params(k) for k in (nb.QUERY_NEURAL_CLOSEST, nb.QUERY_NEURAL_ANY)}
    start = time.perf_counter()
    due = start
    for (sc, kind, off, n, host), gap in zip(reqs, gaps):
        due += gap
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            if due - now > 2e-4:
                time.sleep(due - now - 1e-4)
        if host:
            if hring + n > host_ring:
                hring = 0
            t = srv.submit(scenes[sc], ps[kind], host_pool[sc][off:off + n], h_hits[hring:hring + n])
            hring += n
        else:
            if ring + n > ring_rays:
                ring = 0
            t = srv.submit_device(scenes[sc], ps[kind], pool[sc].ptr + off * 32, n, d_hits.ptr + ring * 48)
            ring += n
        with cond:
            tickets.append((t, n, host))
            cond.notify()
    with cond:
        done[0] = True
        cond.notify()
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([(i["complete_ns"] - i["submit_ns"]) * 1e-6 for i, _, _ in infos])
    is_host = np.array([h for _, _, h in infos])
    t_first = min(i["submit_ns"] for i, _, _ in infos)
    t_last = max(i["complete_ns"] for i, _, _ in infos)
    rays = sum(n for _, n, _ in infos)
    mrays = rays / ((t_last - t_first) * 1e-9) / 1e6
    big.close()
    for h in pinned + [pinned_hits]:
        h.free()
    by_kind = {}
    for name, m in (("device_requests", ~is_host), ("host_requests", is_host)):
        if m.any():
            by_kind[name] = {"requests": int(m.sum()), "p50_ms": round(float(np.percentile(lat[m], 50)), 3),
                             "p99_ms": round(float(np.percentile(lat[m], 99)), 3)}
    return {
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3), **by_kind, "mrays_per_s": round(mrays, 1), "requests": len(infos),
        "rays": int(rays), "offered_requests_per_s": round(rate, 1), "load": load,
        "mix_capacity_mrays_per_s": round(capacity / 1e6, 1),
        "requests_per_launch": round(st["n_batched_requests"] / max(1, st["n_launches"]), 2),
        "workload": "3 resident scenes (81,920 / 2.17M / 8.4M triangles; N-BVH 256 / 4096 / 16384), FP16, "
                    "Poisson arrivals, lognormal sizes median 64k in [4k, 2M], 70% closest / 30% shadow any-hit, "
                    f"{100 - round(100 * host_fraction)}% device-resident / {round(100 * host_fraction)}% host-resident "
                    f"rays, {n_streams} streams, launches of at most {max_batch_rays} rays",
    }


def precision_report(nb, hip, ctx, scene, oracle, module, d_rays, d_hits, stream, W, H, every_row=16, steps=10):
    """On one scene and its 1024x1024 primary rays: per precision the fused query's time, and on
    every 16th image row its hit/miss decisions against the FP64 oracle query of the same rays;
    on those rows' FP64 segments, the per-output errors of the network kernels."""
    cfg, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    net = oracle.Net(cfg, tables, mlp, bmin, bmax)
    rays = d_rays.download(nb.RAY_DTYPE)
    rows = np.arange(0, H, every_row)
    idx = (rows[:, None] * W + np.arange(W)[None, :]).reshape(-1)
    e64, seglog = net.neural_query(scene.nbvh(), rays[idx], mode=oracle.FP64, log_segments=True)
    x = net.encode(rays[idx][seglog["ray"]], seglog["mint"], seglog["maxt"], mode=oracle.FP32)
    z64 = seglog["logit"].astype(np.float64)
    out = {"sample": f"every {every_row}th row ({len(idx)} rays, {len(x)} FP64 segments)",
           "hit_fraction_fp64": sig4((nb.hit_kind(e64) != 0).mean())}
    for name, prec in (("fp32", nb.PREC_FP32), ("fp16", nb.PREC_FP16), ("bf16", nb.PREC_BF16), ("fp8", nb.PREC_FP8)):
        p = module.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=prec, image_width=W)
        module.query_device(p, d_rays.ptr, W * H, d_hits.ptr, stream=stream.handle)
        ev = [hip.Event() for _ in range(steps + 1)]
        ev[0].record(stream)
        for i in range(steps):
            module.query_device(p, d_rays.ptr, W * H, d_hits.ptr, stream=stream.handle)
            ev[i + 1].record(stream)
        stream.synchronize()
        ms = float(np.mean([ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)]))
        hits = d_hits.download(nb.HIT_DTYPE)
        row = {"ms": round(ms, 4), "mrays_per_s": round(W * H / (ms * 1e-3) / 1e6, 1),
               "query_kind_mismatch_vs_fp64": sig4(np.mean(nb.hit_kind(hits[idx]) != nb.hit_kind(e64)))}
        if prec != nb.PREC_FP32 and len(x):
            _, z = module.network_inference(x, precision=prec)
            row["network_errors_vs_fp64"] = output_errors(z, z64)
        out[name] = row
    for name in ("fp16", "bf16", "fp8"):
        out[name]["speedup_vs_fp32"] = round(out["fp32"]["ms"] / out[name]["ms"], 3)
    return out


def trained_network_bench(nb, hip, ctx, oracle, W=1024, H=1024, steps=1500, batch=1 << 16):
    """The headline network trained the reference's way (tree-cut learning of the N-BVH from its
    root with the training step, nbvh_learner_* + nbvh_trainer_*), then the precision report on
    that network and learned N-BVH."""
    v, t = nb.icosphere(6, seed=7, amplitude=0.1)
    scene = nb.Scene(ctx, v, t)
    scene.init_network(nb.network_config(), seed=1, table_range=1e-4)
    lr = nb.Learner(ctx, scene, batch_size=batch, max_training_steps=steps, splitting_interval=25,
                    last_split_iteration=int(steps * 0.75))
    tr = nb.Trainer(ctx, scene, max_batch=batch)
    losses = []
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        losses.append(lr.step(tr))
    el = time.perf_counter() - t0
    tr.commit()
    st = lr.state()
    tr.close()
    lr.close()
    log(f"trained network: {steps} learner steps in {el:.2f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}, "
        f"N-BVH {st['n_nodes']} nodes")
    module = nb.NeuralBVHModule(ctx, scene)
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)
    stream = hip.Stream()
    d_rays, d_hits = hip.Buffer(W * H * 32), hip.Buffer(W * H * 48)
    ctx.raygen(cam, W, H, d_rays.ptr, stream=stream.handle)
    stream.synchronize()
    rep = precision_report(nb, hip, ctx, scene, oracle, module, d_rays, d_hits, stream, W, H)
    stream.destroy()
    scene.close()
    return {"workload": f"headline network and icosphere, trained from the N-BVH root by {steps} learner steps of "
                        f"{batch} samples (tree-cut learning + training step), {W}x{H} primary rays",
            "learner_ms_per_step": round(el / steps * 1e3, 3), "loss_first": sig4(losses[0]),
            "loss_last": sig4(np.mean(losses[-20:])), "nbvh_nodes": st["n_nodes"],
            "sample_failures": st["sample_failures"], "precisions": rep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gather-ceiling", action="store_true", help="skip the random-gather line-rate measurement")
    ap.add_argument("--no-train", action="store_true", help="skip the training-step measurement")
    ap.add_argument("--no-fp32", action="store_true", help="skip timing the FP32 plain path")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (continuous batching) run")
    ap.add_argument("--no-wavefront", action="store_true", help="skip the 4M mixed-ray wavefront run")
    ap.add_argument("--no-path-traced", action="store_true", help="skip the 1920x1080 4-bounce precision run")
    ap.add_argument("--serving-seconds", type=float, default=4.0)
    ap.add_argument("--serving-load", type=float, default=0.7, help="offered load / capacity of the mix")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of MIB MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the neural query is not sharded across GPUs here (rays are "
              "independent, so a batch would be split evenly, one share per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import nbvh_amd as nb
    from nbvh_amd import hip

    W, H = args.width, args.height
    n_rays = W * H
    prec = {"fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp32": nb.PREC_FP32}[args.precision]

    log("building scene")
    v, t = nb.icosphere(6, seed=7, amplitude=0.1)
    ctx = nb.Context()
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    cfg = nb.network_config()
    scene.init_network(cfg, seed=1, table_range=1e-4)
    module = nb.NeuralBVHModule(ctx, scene)
    cam = nb.camera_look_at((0, 0, 3), (0, 0, 0), (0, 1, 0), 45.0, W, H)

    stream = hip.Stream()
    d_rays = hip.Buffer(n_rays * 32)
    d_hits = hip.Buffer(n_rays * 48)
    ctx.raygen(cam, W, H, d_rays.ptr, stream=stream.handle)
    stream.synchronize()

    def timed(precision, steps, warmup):
        p = module.params(kind=nb.QUERY_NEURAL_CLOSEST, precision=precision, image_width=W)
        for _ in range(warmup):
            module.query_device(p, d_rays.ptr, n_rays, d_hits.ptr, stream=stream.handle)
        stream.synchronize()
        hip.synchronize()
        evs = [hip.Event() for _ in range(steps + 1)]
        t0 = time.perf_counter()
        evs[0].record(stream)
        for i in range(steps):
            module.query_device(p, d_rays.ptr, n_rays, d_hits.ptr, stream=stream.handle)
            evs[i + 1].record(stream)
        stream.synchronize()
        hip.synchronize()
        wall = time.perf_counter() - t0
        per = np.array([evs[i].elapsed_ms(evs[i + 1]) for i in range(steps)])
        return per, wall

    log(f"timing {args.precision}: warmup {args.warmup}, steps {args.steps}")
    per, wall = timed(prec, args.steps, args.warmup)
    hits = d_hits.download(nb.HIT_DTYPE)
    segs = int(nb.hit_segments(hits).astype(np.int64).sum())
    ms = float(per.mean())
    value = n_rays / (ms * 1e-3) / 1e6
    log(f"{args.precision}: {ms:.3f} ms/step, {value:.1f} Mrays/s, {segs} segments")

    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic: procedural displaced icosphere, seeded hash tables and MLP weights",
        "config": {"workload": f"neural closest-hit query, {W}x{H} primary rays, icosphere 81920 tris, "
                               "N-BVH 256 leaves, hash grid 16x2 2^19 base 16 scale 1.5, MLP 96-64x4-8",
                   "rays": n_rays, "segments": segs, "precision": args.precision},
        "latency_ms": {"p50": round(float(np.percentile(per, 50)), 4),
                       "p99": round(float(np.percentile(per, 99)), 4)},
        "wall_s_timed": round(wall, 4),
    }

    # Roofline of the dominant kernel (the query kernel is the only kernel of a step besides a
    # one-thread counter reset). Algorithmic bytes: hash-grid gathers 3 points x 16 levels x
    # 8 corners x 2 features x table bytes per segment, plus 32 B ray in + 48 B hit out per ray.
    tb = 4 if prec == nb.PREC_FP32 else 2
    alg_bytes = segs * 3 * 16 * 8 * 2 * tb + n_rays * (32 + 48)
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    result["roofline"] = {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(gbs / HBM_PEAK_GBS, 5),
                          "traffic": measured_traffic(args.precision, W, H),
                          "bytes_per_launch": alg_bytes,
                          "mfma_tflops": round(segs * FLOP_PER_SEGMENT / (ms * 1e-3) / 1e12, 3)}

    if not args.no_fp32 and prec != nb.PREC_FP32:
        per32, _ = timed(nb.PREC_FP32, max(3, args.steps // 4), 1)
        ms32 = float(per32.mean())
        v32 = n_rays / (ms32 * 1e-3) / 1e6
        result["fp32_plain"] = {"value": round(v32, 3), "ms_per_step": round(ms32, 4)}
        result["speedup_vs_fp32_plain"] = round(value / v32, 3)
        log(f"fp32 plain: {ms32:.3f} ms/step, {v32:.1f} Mrays/s -> speedup {value / v32:.2f}x")

    # Visibility agreement with the FP64 oracle on a sample of the workload's segments (the
    # oracle's segment log of every 16th row), evaluated by the same network kernel.
    try:
        import oracle
        cfg2, tables, mlp = scene.network()
        bmin, bmax = scene.bbox()
        net = oracle.Net(cfg2, tables, mlp, bmin, bmax)
        rays = d_rays.download(nb.RAY_DTYPE)
        sample = rays.reshape(H, W)[::16].reshape(-1)
        _, seglog = net.neural_query(scene.nbvh(), sample, mode=oracle.FP64, log_segments=True)
        x = net.encode(sample[seglog["ray"]], seglog["mint"], seglog["maxt"], mode=oracle.FP32)
        _, z = module.network_inference(x, precision=prec)
        flips = float(np.mean((z[:, 0] > 0) != (seglog["logit"][:, 0] > 0)))
        result["vis_flip_rate"] = flips
        result["vis_flip_sample_segments"] = int(len(x))
        log(f"visibility flips vs FP64 oracle: {flips:.5f} over {len(x)} segments")
        # The fused query's own hit/miss decisions against the FP64 oracle query, per precision,
        # with BF16 and FP8 timed beside FP16 and FP32.
        rep = precision_report(nb, hip, ctx, scene, oracle, module, d_rays, d_hits, stream, W, H)
        result["precisions"] = rep
        result["fp16_query_kind_mismatch"] = rep["fp16"]["query_kind_mismatch_vs_fp64"]
        log("precisions: " + ", ".join(f"{k} {rep[k]['mrays_per_s']} Mrays/s mismatch "
                                       f"{rep[k]['query_kind_mismatch_vs_fp64']}" for k in ("fp32", "fp16", "bf16", "fp8")))
        if not args.no_train:
            result["trained_network"] = trained_network_bench(nb, hip, ctx, oracle)

        if not args.no_cpu_baseline:
            result["plain_triangles_512"] = plain_triangles(nb, hip, ctx, scene, oracle)
        if prec != nb.PREC_FP32:
            result["chess_nbvh"] = chess_bench(nb, hip, ctx, oracle)
        if not args.no_gather_ceiling:
            result["mlp_mfma"] = mlp_mfma_bench(nb, hip, ctx)
            result["gather_ceiling"] = gather_ceiling()
            # The query's memory-side lines (its profiled FETCH_SIZE + WRITE_SIZE traffic / 128 B;
            # the hit writes are ~1 % of it) against the random-line rate measured for a table of
            # the FP16 tables' size (26 MB -> 32 MiB).
            traffic = result["roofline"]["traffic"]
            if traffic and prec == nb.PREC_FP16 and (W, H) == (1024, 1024):
                achieved = traffic / 128 / (ms * 1e-3) / 1e9
                ceiling = result["gather_ceiling"]["32MiB_glines_per_s"]
                result["roofline"]["line_rate"] = {"achieved_glines_per_s": round(achieved, 2),
                                                   "ceiling_glines_per_s": ceiling,
                                                   "frac": round(achieved / ceiling, 3)}
                log(f"query memory-side line rate {achieved:.1f} G lines/s = {achieved / ceiling:.3f} of the "
                    f"{ceiling} G lines/s random-line ceiling of a 32 MiB table")
        if not args.no_train:
            result["training"] = train_bench(nb, hip, ctx, oracle)

        if not args.no_path_traced and prec != nb.PREC_FP32:
            pt = path_traced_bench(nb, hip, ctx, scene, oracle)
            result["path_traced"] = pt
            result["hybrid_path_traced"] = hybrid_path_traced_bench(nb, hip, ctx, scene)
            result["path_trace_driver"] = path_trace_driver_bench(nb, hip, ctx, scene)

        if not args.no_cpu_baseline:
            # CPU baseline: the oracle's FP32 path (same traversal + hash grid + FP32 MLP), all
            # host threads OpenMP-parallel, on the full frame; repeated to >= 10 s of CPU time.
            nodes = scene.nbvh()
            net.neural_query(nodes, rays[:4096], mode=oracle.FP32)
            reps, t0 = 0, time.perf_counter()
            while True:
                net.neural_query(nodes, rays, mode=oracle.FP32)
                reps += 1
                el = time.perf_counter() - t0
                if (reps >= 3 and el * oracle.num_threads() >= 10.0) or el > 30.0:
                    break
            cpu = reps * n_rays / el / 1e6
            result["cpu_baseline"] = {"value": round(cpu, 4), "unit": "Mrays/s", "cores": oracle.num_threads(),
                                      "kind": "port",
                                      "sample": f"{reps} x full {W}x{H} frame, oracle FP32 path, {el:.1f} s wall"}
            log(f"cpu baseline: {cpu:.3f} Mrays/s on {oracle.num_threads()} threads")
            # One thread, on every 8th row of the frame (a few seconds of CPU work).
            threads = oracle.num_threads()
            oracle.set_num_threads(1)
            sub = rays.reshape(H, W)[::8].reshape(-1)
            t0 = time.perf_counter()
            net.neural_query(nodes, sub, mode=oracle.FP32)
            el1 = time.perf_counter() - t0
            oracle.set_num_threads(threads)
            result["cpu_baseline_1_thread"] = {"value": round(len(sub) / el1 / 1e6, 4), "unit": "Mrays/s",
                                               "sample": f"every 8th row of the {W}x{H} frame ({len(sub)} rays), "
                                                         f"{el1:.1f} s"}
            log(f"cpu baseline, 1 thread: {len(sub) / el1 / 1e6:.4f} Mrays/s")
    except ImportError as e:
        log(f"oracle unavailable ({e}); no flip rate / cpu baseline")

    if not args.no_wavefront and prec != nb.PREC_FP32:
        wf = wavefront_bench(nb, hip, ctx, max(5, args.steps // 2))
        result["wavefront"] = wf
        log(f"wavefront: fp16 {wf['fp16']['mrays_per_s']} Mrays/s coherence-ordered (array order "
            f"{wf['fp16_array_order']['mrays_per_s']}, graph in array order {wf['fp16_graph']['mrays_per_s']}, "
            f"Morton-sorted rounds {wf['fp16_wavefront_rounds']['mrays_per_s']}), fp32 {wf['fp32']['mrays_per_s']} "
            f"(array order {wf['fp32_array_order']['mrays_per_s']}, rounds "
            f"{wf['fp32_wavefront_rounds']['mrays_per_s']}), triangles {wf['triangle_fallback']['mrays_per_s']}")

    if not args.no_serving and prec != nb.PREC_FP32:
        sv = serving_bench(nb, hip, ctx, scene, args.serving_seconds, args.serving_load)
        result["serving"] = sv
        result["p99_ms"] = sv["p99_ms"]
        log(f"serving: p50 {sv['p50_ms']} ms, p99 {sv['p99_ms']} ms, {sv['mrays_per_s']} Mrays/s, "
            f"{sv['requests']} requests, {sv['requests_per_launch']} requests per launch; "
            + ", ".join(f"{k}: {v['requests']} p50 {v['p50_ms']} p99 {v['p99_ms']} ms" for k, v in sv.items()
                        if k.endswith("_requests")))

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = hip.Buffer(nbytes), hip.Buffer(nbytes)
        src.zero()
        hip.synchronize()
        hip.check(hip._runtime().hipMemcpy(dst.ptr, src.ptr, nbytes, hip.D2D), "hipMemcpy D2D")
        hip.synchronize()
        result["calibration_copy_bytes"] = nbytes
        log(f"calibration copy: {nbytes} bytes device to device")

    print(json.dumps(result), flush=True)
    for sc in _EXTRA_SCENES.values():
        sc.close()
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()
