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


def serving_bench(nb, hip, ctx, scene_a, seconds, load, seed=11):
    """Continuous-batching serving run over three resident scenes.

    Scenes: the 81,920-triangle icosphere (N-BVH 256), a field of 64 geodesic spheres (2.17M
    triangles, N-BVH 4096) and one geodesic sphere of frequency 648 (8.4M triangles, N-BVH
    16384), each with its own seeded FP16 network of the headline shape. Requests arrive as a
    Poisson process; sizes are lognormal (median 65,536 rays, sigma 1) clamped to [4,096, 2M];
    70 % closest-hit, 30 % any-hit (shadow) neural queries; the scene is drawn uniformly. Rays
    are windows of a 2048x1024 camera frame of the request's scene, resident in HBM; hits go to
    a device ring (zero-copy device requests). The arrival rate is set to `load` times the
    measured single-launch capacity of the mix. Latency is submit -> hits complete, from the
    server's own steady-clock timestamps."""
    import threading
    rng = np.random.default_rng(seed)
    log("serving: building the 2.17M- and 8.4M-triangle scenes")
    t0 = time.perf_counter()
    scenes = [scene_a]
    for verts_tris, leaves, net_seed in ((nb.sphere_field(64, seed=5), 4096, 2),
                                         (nb.geosphere(648, seed=9, amplitude=0.1), 16384, 3)):
        sc = nb.Scene(ctx, *verts_tris)
        sc.build_nbvh(leaves)
        sc.init_network(nb.network_config(), seed=net_seed, table_range=1e-4)
        scenes.append(sc)
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
    ring_rays = 16 << 20
    d_hits = hip.Buffer(ring_rays * 48)

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
        reqs.append((sc, kind, off, n))
    log(f"serving: capacity of the mix {capacity / 1e6:.0f} Mrays/s, mean request {mean_size:.0f} rays, "
        f"{rate:.0f} requests/s for {seconds:.1f} s ({n_req} requests, load {load})")

    srv = nb.Server(ctx, n_streams=4, max_batch_rays=1 << 21)
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
                t, n = tickets[i]
            _, info = srv.wait(t)
            infos.append((info, n))
            i += 1

    th = threading.Thread(target=collector)
    th.start()
    ring = 0
    ps = {k: params(k) for k in (nb.QUERY_NEURAL_CLOSEST, nb.QUERY_NEURAL_ANY)}
    start = time.perf_counter()
    due = start
    for (sc, kind, off, n), gap in zip(reqs, gaps):
        due += gap
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            if due - now > 2e-4:
                time.sleep(due - now - 1e-4)
        if ring + n > ring_rays:
            ring = 0
        t = srv.submit_device(scenes[sc], ps[kind], pool[sc].ptr + off * 32, n, d_hits.ptr + ring * 48)
        ring += n
        with cond:
            tickets.append((t, n))
            cond.notify()
    with cond:
        done[0] = True
        cond.notify()
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([(i["complete_ns"] - i["submit_ns"]) * 1e-6 for i, _ in infos])
    t_first = min(i["submit_ns"] for i, _ in infos)
    t_last = max(i["complete_ns"] for i, _ in infos)
    rays = sum(n for _, n in infos)
    mrays = rays / ((t_last - t_first) * 1e-9) / 1e6
    for sc in scenes[1:]:
        sc.close()
    return {
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3), "mrays_per_s": round(mrays, 1), "requests": len(infos),
        "rays": int(rays), "offered_requests_per_s": round(rate, 1), "load": load,
        "mix_capacity_mrays_per_s": round(capacity / 1e6, 1),
        "requests_per_launch": round(st["n_batched_requests"] / max(1, st["n_launches"]), 2),
        "workload": "3 resident scenes (81,920 / 2.17M / 8.4M triangles; N-BVH 256 / 4096 / 16384), FP16, "
                    "Poisson arrivals, lognormal sizes median 64k in [4k, 2M], 70% closest / 30% shadow any-hit, "
                    "device-resident rays, 4 streams",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip timing the FP32 plain path")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (continuous batching) run")
    ap.add_argument("--serving-seconds", type=float, default=4.0)
    ap.add_argument("--serving-load", type=float, default=0.7, help="offered load / capacity of the mix")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of MIB MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the neural query is not sharded across GPUs in this build "
              "(rays are independent; see DESIGN.md for how the batch would shard). Run with --gpus 1.",
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
                if el * oracle.num_threads() >= 10.0 or el > 30.0:
                    break
            cpu = reps * n_rays / el / 1e6
            result["cpu_baseline"] = {"value": round(cpu, 4), "unit": "Mrays/s", "cores": oracle.num_threads(),
                                      "kind": "port",
                                      "sample": f"{reps} x full {W}x{H} frame, oracle FP32 path, {el:.1f} s wall"}
            log(f"cpu baseline: {cpu:.3f} Mrays/s on {oracle.num_threads()} threads")
    except ImportError as e:
        log(f"oracle unavailable ({e}); no flip rate / cpu baseline")

    if not args.no_serving and prec != nb.PREC_FP32:
        sv = serving_bench(nb, hip, ctx, scene, args.serving_seconds, args.serving_load)
        result["serving"] = sv
        result["p99_ms"] = sv["p99_ms"]
        log(f"serving: p50 {sv['p50_ms']} ms, p99 {sv['p99_ms']} ms, {sv['mrays_per_s']} Mrays/s, "
            f"{sv['requests']} requests, {sv['requests_per_launch']} requests per launch")

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
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()
