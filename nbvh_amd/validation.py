"""Validation-mode batch harness: the reference's `ntwr <scene> true <config.json>...` (src/main.cpp:32-57)
over this library.

For each configuration file, in order (RenderAppConfigs / NeuralBVHRenderer::load_next_config,
src/cuda/neural/neural_bvh_renderer_init.cu:146-245), its "neural_bvh_renderer.validation_mode":
  Training   tree-cut learning of the N-BVH and training of the network for max_training_steps
             steps (nbvh_learner_* + nbvh_trainer_*, the loop of neural_bvh_renderer.cu:141-297),
             then results/<base>_<variant>_learning_spp_0.csv (per-step time and loss),
             nbvhs/<base>.nbvh and network_models/<base>_network.msgpack
             (process_validation_config_end / save_performance_metrics / save_neural_bvh /
             save_network_model, neural_bvh_renderer_init.cu:247-383);
  Inference  loads nbvhs/<base>.nbvh and network_models/<base>_network.msgpack, renders
             max_accumulated_spp 1-spp frames of diffuse path tracing with the neural object, then
             results/<base>_<variant>_spp_<spp>.exr and results/<base>_<variant>_inference_spp_<spp>.csv
             (per-frame time);
  Full       Training, then Inference on the trained state.
Every rendering is compared with the triangle rendering of the same object at the same samples
(ignore_neural_blas) by FLIP (compute_and_display_flip_error, neural_bvh_renderer.cu:40-92):
results/<base>_<variant>_flip.json holds the mean error and the reference image goes to
results/<base>_reference_spp_<spp>.exr.

The geometry: the reference loads a glTF scene given on its command line; its scenes are not
shipped with it, so the object here is a procedural mesh (--scene icosphere:SUBDIV:SEED:AMPLITUDE or
sphere_field:COUNT:SEED). Settings read from each configuration (missing ones keep the reference's
defaults): batch_size, max_training_steps, max_accumulated_spp, base/variant names,
neural_bvh_min_depth, tree_cut_learning, bvh_split_scheduler, bvh_constants (sampler, survival,
num_lods_learned, rendered_lod), neural_module.max_depth, pt_output_constants visibility
thresholds, path_module_constants (rr_start_depth, normal_epsilon_offset, apply_russian_roulette),
optimizer, segment_encoding.hashgrid_options, network, bvh_module_constants (points,
distribution), cuda_scene.fb_size and skylike_background.

  python -m nbvh_amd.validation CONFIG.json [...] --out DIR [--scene icosphere:6:7:0.1]
         [--precision fp16|bf16|fp32] [--width W --height H] [--spp N] [--steps N]
"""
import argparse
import csv
import json
import os
import struct
import sys
import time

import numpy as np

import nbvh_amd as nb
from nbvh_amd import _ffi

PRECISIONS = {"fp32": nb.PREC_FP32, "fp16": nb.PREC_FP16, "bf16": nb.PREC_BF16, "fp8": nb.PREC_FP8}


# ---------------------------------------------------------------------------------------------
# OpenEXR (scanline, uncompressed, FLOAT R G B), the format of the reference's save_framebuffer.
# ---------------------------------------------------------------------------------------------
def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def write_exr(path, rgb):
    """rgb: float [h, w, 3] -> an OpenEXR file (single part, scanline, NO_COMPRESSION, FLOAT)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    chans = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1) for c in ("B", "G", "R")) + b"\0"
    header = b"".join([
        _attr("channels", "chlist", chans),
        _attr("compression", "compression", b"\0"),
        _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)),
        _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)),
        _attr("lineOrder", "lineOrder", b"\0"),
        _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)),
        _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ]) + b"\0"
    prefix = struct.pack("<I", 20000630) + struct.pack("<I", 2) + header
    line_bytes = 8 + 3 * 4 * w
    first = len(prefix) + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * line_bytes for y in range(h)])
    with open(path, "wb") as f:
        f.write(prefix)
        f.write(offsets)
        for y in range(h):
            row = rgb[y]
            f.write(struct.pack("<ii", y, 3 * 4 * w))
            f.write(np.ascontiguousarray(row[:, 2]).tobytes())
            f.write(np.ascontiguousarray(row[:, 1]).tobytes())
            f.write(np.ascontiguousarray(row[:, 0]).tobytes())


def read_exr(path):
    """Reads the files write_exr writes (scanline, uncompressed, FLOAT channels) -> [h, w, 3] (R, G, B)."""
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<II", data, 0)
    if magic != 20000630 or (version & 0xFF) != 2:
        raise ValueError("not a single-part scanline OpenEXR file")
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        size = struct.unpack_from("<i", data, e2 + 1)[0]
        attrs[name] = data[e2 + 5:e2 + 5 + size]
        pos = e2 + 5 + size
    pos += 1
    if attrs["compression"] != b"\0":
        raise ValueError("only uncompressed files are read")
    names, p, cl = [], 0, attrs["channels"]
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        names.append(cl[p:e].decode())
        if struct.unpack_from("<i", cl, e + 1)[0] != 2:
            raise ValueError("only FLOAT channels are read")
        p = e + 17
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    out = np.zeros((h, w, 3), np.float32)
    for off in offsets:
        y = struct.unpack_from("<i", data, off)[0] - y0
        px = np.frombuffer(data, "<f4", len(names) * w, off + 8).reshape(len(names), w)
        for c, name in enumerate(names):
            if name in ("R", "G", "B"):
                out[y, :, "RGB".index(name)] = px[c]
    return out


# ---------------------------------------------------------------------------------------------
# Configuration
# ---------------------------------------------------------------------------------------------
def settings(config):
    """The harness settings of one reference configuration (defaults where a field is missing)."""
    r = config.get("neural_bvh_renderer", {})
    nm = r.get("neural_module", {})
    bc = r.get("bvh_constants", {})
    sched = r.get("bvh_split_scheduler", {})
    pt = nm.get("pt_output_constants", {})
    pm = nm.get("path_module_constants", {})
    opt = nm.get("optimizer", {})
    grid = nm.get("segment_encoding", {}).get("hashgrid_options", {})
    net = nm.get("network", {})
    bm = nm.get("bvh_module_constants", {})
    scene = config.get("cuda_scene", {})
    learner = nb.Learner.defaults()
    learner.update(
        batch_size=int(r.get("batch_size", 1 << 18)),
        log2_bvh_max_traversals=int(bc.get("log2_bvh_max_traversals", 0)),
        sampling_bbox_scaling=float(bc.get("sampling_bbox_scaling", 1.5)),
        fixed_node_dilation=float(bc.get("fixed_node_dilation", 1e-5)),
        node_extent_dilation=float(bc.get("node_extent_dilation", 1e-4)),
        learning_mint_epsilon=float(bc.get("learning_mint_epsilon", 1e-3)),
        learning_maxt_epsilon=float(bc.get("learning_maxt_epsilon", 1e-3)),
        max_training_bounces=int(bc.get("max_training_bounces", 1)),
        training_node_survival=int(bool(bc.get("training_node_survival", True))),
        survival_probability_scale=float(bc.get("survival_probability_scale", 1.0)),
        min_rr_survival_prob=float(bc.get("min_rr_survival_prob", 0.005)),
        tree_cut_learning=int(bool(r.get("tree_cut_learning", True))),
        neural_bvh_min_depth=int(r.get("neural_bvh_min_depth", 0)),
        num_lods_learned=int(bc.get("num_lods_learned", 1)),
        max_training_steps=int(r.get("max_training_steps", 10000)),
        splitting_interval=int(sched.get("splitting_interval", 50)),
        last_split_iteration=int(sched.get("last_split_iteration", 2000)),
        start_split_num=int(sched.get("start_split_num", 1)),
        split_scaling=float(sched.get("split_scaling", 2.0)),
        split_interval_scaling=float(sched.get("split_interval_scaling", 2.0)),
    )
    dist = {"EquidistantInterior": 0, "EquidistantExterior": 1}.get(bm.get("segment_distribution",
                                                                            "EquidistantExterior"))
    if dist is None:
        # Uniform is a query option here (nbvh_query with segment_seed); the harness's learner and
        # path tracer refuse it.
        raise ValueError(f"segment_distribution {bm.get('segment_distribution')} is not supported by the harness")
    depth_mode = {"LocalT": 0, "WorldT": 1}.get(pt.get("depth_learning_mode", "LocalT"))
    if depth_mode is None:
        raise ValueError(f"depth_learning_mode {pt.get('depth_learning_mode')} is not supported")
    return dict(
        mode=r.get("validation_mode", "Ignore"),
        base=r.get("base_config_name", "config"), variant=r.get("variant_config_name", "variant"),
        steps=int(r.get("max_training_steps", 10000)), spp=int(r.get("max_accumulated_spp", 1)),
        learner=learner,
        trainer=dict(learning_rate=float(opt.get("learning_rate", 1e-2)), beta1=float(opt.get("beta1", 0.9)),
                     beta2=float(opt.get("beta2", 0.999)), epsilon=float(opt.get("epsilon", 1e-8)),
                     l2_reg=float(opt.get("l2_reg", 1e-8)), decay_start=int(opt.get("decay_start", 10000)),
                     decay_interval=int(opt.get("decay_interval", 2000)), decay_base=float(opt.get("decay_base", 0.33))),
        network=nb.network_config(n_levels=int(grid.get("n_levels", 8)),
                                  n_features_per_level=int(grid.get("n_features_per_level", 4)),
                                  log2_hashmap_size=int(grid.get("log2_hashmap_size", 18)),
                                  base_resolution=int(grid.get("base_resolution", 8)),
                                  per_level_scale=float(grid.get("per_level_scale", 2.0)),
                                  n_points=int(bm.get("input_segment_points", 3)), segment_distribution=dist,
                                  n_neurons=int(net.get("n_neurons", 64)),
                                  n_hidden_layers=int(net.get("n_hidden_layers", 4)), depth_mode=depth_mode,
                                  depth_linear_activation=bool(pt.get("output_depth_linear_activation", False)),
                                  interpolate_depth=bool(pt.get("interpolate_depth", False)),
                                  depth_interpolation_eps=float(pt.get("depth_interpolation_eps", 1e-3))),
        max_depth=int(nm.get("max_depth", 3)),
        primary_threshold=float(pt.get("primary_visibility_threshold", 0.5)),
        secondary_threshold=float(pt.get("secondary_visibility_threshold", 0.5)),
        rr_start_depth=int(pm.get("rr_start_depth", 2)) if pm.get("apply_russian_roulette", True) else 0,
        normal_offset=float(pm.get("normal_epsilon_offset", 1e-4)),
        lod=int(bc.get("rendered_lod", 0)),
        fb_size=tuple(scene.get("fb_size", (1920, 1080))),
        sky=bool(scene.get("scene_constants", {}).get("skylike_background", True)),
    )


def make_mesh(spec):
    kind, *args = spec.split(":")
    if kind == "icosphere":
        sub, seed, amp = (int(args[0]) if args else 6), (int(args[1]) if len(args) > 1 else 7), \
            (float(args[2]) if len(args) > 2 else 0.1)
        return nb.icosphere(sub, seed=seed, amplitude=amp)
    if kind == "sphere_field":
        return nb.sphere_field(int(args[0]) if args else 64, seed=int(args[1]) if len(args) > 1 else 5)
    raise ValueError(f"unknown scene {spec}")


# ---------------------------------------------------------------------------------------------
# The validation loop
# ---------------------------------------------------------------------------------------------
def train(ctx, scene, s, out, log):
    lr = nb.Learner(ctx, scene, **s["learner"])
    tr = nb.Trainer(ctx, scene, max_batch=s["learner"]["batch_size"], **s["trainer"])
    times, losses = [], []
    for step in range(s["steps"]):
        t0 = time.perf_counter()
        losses.append(lr.step(tr))
        times.append((time.perf_counter() - t0) * 1e3)
        if step % 500 == 0 or step == s["steps"] - 1:
            st = lr.state()
            log(f"  step {step}: loss {losses[-1]:.5f}, {st['n_nodes']} N-BVH nodes, LoD {st['current_learned_lod']}")
    tr.commit()
    state = lr.state()
    tr.close()
    lr.close()
    os.makedirs(os.path.join(out, "results"), exist_ok=True)
    with open(os.path.join(out, "results", f"{s['base']}_{s['variant']}_learning_spp_0.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["frame index", "frame time (ms)", " loss"])
        for i, (t, l) in enumerate(zip(times, losses)):
            w.writerow([i, f"{t:.6f}", f"{l:.9g}"])
    os.makedirs(os.path.join(out, "nbvhs"), exist_ok=True)
    with open(os.path.join(out, "nbvhs", f"{s['base']}.nbvh"), "wb") as f:
        f.write(nb.nbvh_serialize(scene.nbvh()))
    write_network(os.path.join(out, "network_models", f"{s['base']}_network.msgpack"), scene)
    return dict(steps=len(times), ms_per_step=float(np.mean(times)) if times else 0.0,
                loss_first=losses[0] if losses else None, loss_last=losses[-1] if losses else None,
                n_nodes=state["n_nodes"], total_splits=state["total_splits"],
                sample_failures=state["sample_failures"])


def write_network(path, scene):
    """The network as the reference's saved model object (n_params, params_binary of the flat
    tiny-cuda-nn parameters), MessagePack-encoded; read back by nbvh_amd.read_tcnn_model."""
    import msgpack
    cfg, tables, mlp = scene.network()
    params = nb.network_to_tcnn_params(cfg, tables, mlp)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(msgpack.packb({"n_params": int(len(params)), "params_binary": params.astype("<f4").tobytes()}))


def render(ctx, scene, s, out, width, height, precision, log):
    mod = nb.NeuralBVHModule(ctx, scene)
    lo, hi = scene.bbox()
    c = (lo + hi) * 0.5
    ext = float(np.max(hi - lo))
    cam = nb.camera_look_at(tuple(c + np.array([0.35, 0.45, 1.0]) * 1.7 * ext), tuple(c), (0, 1, 0), 45.0, width,
                            height)
    from nbvh_amd import hip
    acc_n, acc_t = hip.Buffer(width * height * 16), hip.Buffer(width * height * 16)
    common = dict(max_depth=s["max_depth"], rr_start_depth=s["rr_start_depth"], normal_offset=s["normal_offset"],
                  background=nb.BACKGROUND_SKY if s["sky"] else nb.BACKGROUND_CONSTANT)
    times = []
    for spp in range(s["spp"]):
        hip.synchronize()
        t0 = time.perf_counter()
        mod.path_trace_device(cam, width, height, acc_n.ptr, precision=precision, sample=spp, lod=s["lod"],
                              visibility_threshold=s["primary_threshold"],
                              secondary_visibility_threshold=s["secondary_threshold"], **common)
        hip.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        mod.path_trace_device(cam, width, height, acc_t.ptr, sample=spp, triangles=True,
                              environment_albedo=(0.9, 0.9, 0.9), **common)
    hip.synchronize()
    neural = acc_n.download(np.float32).reshape(height, width, 4)
    tri = acc_t.download(np.float32).reshape(height, width, 4)
    err, _ = nb.flip(ctx, tri, neural)
    res = os.path.join(out, "results")
    os.makedirs(res, exist_ok=True)
    write_exr(os.path.join(res, f"{s['base']}_{s['variant']}_spp_{s['spp']}.exr"), neural[..., :3] / neural[..., 3:4])
    write_exr(os.path.join(res, f"{s['base']}_reference_spp_{s['spp']}.exr"), tri[..., :3] / tri[..., 3:4])
    with open(os.path.join(res, f"{s['base']}_{s['variant']}_inference_spp_{s['spp']}.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["frame index", "frame time (ms)"])
        for i, t in enumerate(times):
            w.writerow([i, f"{t:.6f}"])
    summary = dict(spp=s["spp"], width=width, height=height, flip=err, ms_per_frame=float(np.mean(times)))
    with open(os.path.join(res, f"{s['base']}_{s['variant']}_flip.json"), "w") as f:
        json.dump(summary, f, indent=1)
    log(f"  rendered {s['spp']} spp at {width}x{height}: FLIP {err:.4f} against the triangle rendering, "
        f"{summary['ms_per_frame']:.2f} ms per frame")
    return summary


def load_trained(ctx, scene, s, out):
    with open(os.path.join(out, "nbvhs", f"{s['base']}.nbvh"), "rb") as f:
        scene.load_neural_bvh(f.read())
    scene.load_tcnn_model(s["network"], os.path.join(out, "network_models", f"{s['base']}_network.msgpack"))


def run(configs, out, scene_spec="icosphere:6:7:0.1", precision="fp16", width=None, height=None, spp=None,
        steps=None, log=print):
    """Runs the configurations in order; returns one result dict per configuration."""
    ctx = nb.Context()
    verts, tris = make_mesh(scene_spec)
    results = []
    try:
        for path in configs:
            with open(path) as f:
                s = settings(json.load(f))
            if spp is not None:
                s["spp"] = spp
            if steps is not None:
                s["steps"] = steps
                s["learner"]["max_training_steps"] = steps
            w, h = width or s["fb_size"][0], height or s["fb_size"][1]
            log(f"{os.path.basename(path)}: {s['mode']} ({s['base']} / {s['variant']})")
            if s["mode"] not in ("Training", "Inference", "Full"):
                log("  skipped: validation_mode is not Training, Inference or Full")
                results.append(dict(config=path, skipped=True))
                continue
            scene = nb.Scene(ctx, verts, tris)
            r = dict(config=path, mode=s["mode"], base=s["base"], variant=s["variant"])
            if s["mode"] in ("Training", "Full"):
                scene.init_network(s["network"], seed=1, table_range=1e-4)
                r["training"] = train(ctx, scene, s, out, log)
            if s["mode"] in ("Inference", "Full"):
                if s["mode"] == "Inference":
                    load_trained(ctx, scene, s, out)
                r["inference"] = render(ctx, scene, s, out, w, h, PRECISIONS[precision], log)
            scene.close()
            results.append(r)
    finally:
        ctx.close()
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("configs", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--scene", default="icosphere:6:7:0.1")
    ap.add_argument("--precision", default="fp16", choices=sorted(PRECISIONS))
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--spp", type=int, help="override max_accumulated_spp")
    ap.add_argument("--steps", type=int, help="override max_training_steps")
    a = ap.parse_args(argv)
    res = run(a.configs, a.out, a.scene, a.precision, a.width, a.height, a.spp, a.steps,
              log=lambda m: print(m, flush=True))
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
