"""GPU trainer (nbvh_trainer_*) against the FP64 training restatement (oracle/train.py).

Bars: the loss and the matrix gradients within 1e-4 (relative, as vector norms) of the FP64
restatement and the table gradients within 1e-3 (the GPU computes grid cells in FP32, the oracle in
FP64, so a point within an ulp of a cell boundary may scatter into the neighbouring cell); one Adam
step equal to the restatement's step on the GPU's own gradients to 1e-6; and training on segments
of the benchmark mesh's N-BVH leaves labelled from its triangles raises the network's held-out
visibility accuracy well above the class prior.
"""
import numpy as np
import pytest

from conftest import bench_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def setup(nb, orc, ctx):
    v, t, nodes, prims, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(nb.network_config(), seed=1, table_range=1e-2)
    cfg, tables, mlp = scene.network()
    lo, hi = scene.bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    rays, mint, maxt = nb.leaf_segments(scene.nbvh(), 420000, seed=0)
    x = net.encode(rays, mint, maxt, mode=orc.FP32)
    mod = nb.NeuralBVHModule(ctx, scene)
    targets = mod.segment_labels(rays, mint, maxt)  # the scene's triangles label every segment
    yield dict(scene=scene, cfg=cfg, tables=tables, mlp=mlp, net=net, x=x, t=targets, mod=mod)
    scene.close()


def _check_gradients(nb, ctx, scene, cfg, tables, mlp, lv, x, t):
    """One gradients-only step against the FP64 restatement. Targets of the L1 outputs (depth,
    normal) are placed 0.25 away from the prediction, so that no element sits at the kink of |d|
    where FP32 and FP64 could take different signs."""
    from oracle import train
    T64, M64 = tables.astype(np.float64), mlp.astype(np.float64)
    y = train.forward(lv, cfg, T64, M64, x.astype(np.float64))[3]
    t = t.astype(np.float64).copy()
    t[:, 1] = y[:, 1] + np.where(y[:, 1] > 0.5, -0.25, 0.25)
    lin = ((y[:, 2:5] * 2.0 - 1.0) / 0.1) * 0.5 + 0.5
    t[:, 2:5] = lin + np.where(lin > 0.5, -0.25, 0.25)
    t = t.astype(np.float32)
    tr = nb.Trainer(ctx, scene, max_batch=4096)
    loss = tr.step(x, t, gradients_only=True)
    gt, gm = tr.get(nb.TRAIN_GRAD_TABLES), tr.get(nb.TRAIN_GRAD_MLP)
    l64, dt64, dm64, _ = train.train_gradients(lv, cfg, T64, M64, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - l64) <= 1e-4 * abs(l64), (loss, l64)
    assert np.linalg.norm(gm - dm64) <= 1e-4 * np.linalg.norm(dm64)
    assert np.linalg.norm(gt - dt64) <= 1e-3 * np.linalg.norm(dt64)
    assert np.count_nonzero(gt) > 1000 and np.abs(dt64).max() > 0
    # Entries no sample touched have an exactly zero gradient (the Adam step skips them).
    assert np.array_equal(gt == 0, dt64 == 0)
    # gradients only: the parameters are untouched
    assert np.array_equal(tr.get(nb.TRAIN_MLP), mlp) and np.array_equal(tr.get(nb.TRAIN_TABLES), tables)
    tr.close()


@pytest.mark.parametrize("n", [512, 4000])
def test_gradients_match_oracle(nb, orc, ctx, setup, n):
    """n = 4000 puts 96,000 corners into the coarsest level's single slice, which the trainer then
    splits across workgroups (their partial slices combined with global atomics); at n = 512 every
    slice is summed by one workgroup."""
    _check_gradients(nb, ctx, setup["scene"], setup["cfg"], setup["tables"], setup["mlp"], setup["net"].levels(),
                     setup["x"][:n], setup["t"][:n])


@pytest.mark.parametrize("n", [1, 2000])
def test_chess_shape_gradients_match_oracle(nb, orc, ctx, setup, n):
    """The (3, 8, 4) shape of the reference's chess configuration (8 levels x 4 features, 2^18
    entries, base 8, scale 2) on the same segments; n = 1 is the smallest batch."""
    cfg = nb.network_config(n_levels=8, n_features_per_level=4, log2_hashmap_size=18, base_resolution=8,
                            per_level_scale=2.0)
    v, t, _, _, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(cfg, seed=4, table_range=1e-2)
    _, tables, mlp = scene.network()
    lo, hi = scene.bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    rays, mint, maxt = nb.leaf_segments(scene.nbvh(), n, seed=2)
    x = net.encode(rays, mint, maxt, mode=orc.FP32)
    targets = nb.NeuralBVHModule(ctx, scene).segment_labels(rays, mint, maxt)
    if n == 1:
        # a single sample touches at most 3 x 8 x 8 entries
        from oracle import train
        tr = nb.Trainer(ctx, scene, max_batch=64)
        tr.step(x, targets, gradients_only=True)
        gt = tr.get(nb.TRAIN_GRAD_TABLES)
        _, dt64, _, _ = train.train_gradients(net.levels(), cfg, tables.astype(np.float64), mlp.astype(np.float64),
                                              x.astype(np.float64), targets.astype(np.float64))
        assert np.array_equal(gt == 0, dt64 == 0) and 0 < np.count_nonzero(gt) <= 3 * 8 * 8 * 4
        assert np.linalg.norm(gt - dt64) <= 1e-3 * np.linalg.norm(dt64)
        tr.close()
    else:
        _check_gradients(nb, ctx, scene, cfg, tables, mlp, net.levels(), x, targets)
    scene.close()


@pytest.mark.parametrize("shape", [(8, 4, 22, 8, 2.0), (16, 2, 21, 16, 1.5)], ids=["chess_2^22", "headline_2^21"])
def test_large_table_gradients_match_oracle(nb, orc, ctx, setup, shape):
    """Hash tables of 2^22 entries at 4 features (the chess shape) and 2^21 at 2 (the headline
    shape): every hashed level is 128 or more slices of 128 KB, so its corner terms are binned
    (count, scan, scatter, one LDS sum per bin) instead of scanned once per slice, while the small
    dense levels keep the slice scans; the gradients still match the FP64 restatement."""
    n_levels, n_features, log2, base, scale = shape
    cfg = nb.network_config(n_levels=n_levels, n_features_per_level=n_features, log2_hashmap_size=log2,
                            base_resolution=base, per_level_scale=scale)
    v, t, _, _, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, t)
    scene.build_nbvh(256)
    scene.init_network(cfg, seed=6, table_range=1e-2)
    _, tables, mlp = scene.network()
    lo, hi = scene.bbox()
    net = orc.Net(cfg, tables, mlp, lo, hi)
    rays, mint, maxt = nb.leaf_segments(scene.nbvh(), 1500, seed=5)
    x = net.encode(rays, mint, maxt, mode=orc.FP32)
    targets = nb.NeuralBVHModule(ctx, scene).segment_labels(rays, mint, maxt)
    lv = net.levels()
    hashed = lv["dense"] == 0
    slices = lv["size"].astype(np.int64) * n_features * 4 // (128 * 1024)
    assert hashed.sum() >= 3 and np.all(lv["size"][hashed] == 1 << log2) and np.all(slices[hashed] >= 128)
    assert np.any(slices[~hashed] < 128)
    _check_gradients(nb, ctx, scene, cfg, tables, mlp, lv, x, targets)
    scene.close()


def test_adam_step_matches_oracle(nb, orc, ctx, setup):
    from oracle import train
    x, t = setup["x"][:2048], setup["t"][:2048]
    tr = nb.Trainer(ctx, setup["scene"], max_batch=4096)
    tr.step(x, t)
    # The gradients the update used stay in the trainer: the restatement's Adam on them.
    gt, gm = tr.get(nb.TRAIN_GRAD_TABLES), tr.get(nb.TRAIN_GRAD_MLP)
    wt, wm = tr.get(nb.TRAIN_TABLES), tr.get(nb.TRAIN_MLP)
    et, _, _, _ = train.adam_step(setup["tables"].astype(np.float64), gt.astype(np.float64), 0.0, 0.0, 0)
    em, _, _, _ = train.adam_step(setup["mlp"].astype(np.float64), gm.astype(np.float64), 0.0, 0.0, 0,
                                  matrix=True)
    assert np.allclose(wm, em, rtol=0, atol=1e-6)
    assert np.allclose(wt, et, rtol=0, atol=1e-6)
    untouched = gt == 0
    assert untouched.sum() > 0 and np.array_equal(wt[untouched], setup["tables"][untouched])
    assert np.abs(wm - setup["mlp"]).max() > 5e-3  # a first Adam step moves by about the learning rate
    tr.close()


def test_adam_with_lr_decay_matches_oracle(nb, orc, ctx, setup):
    """Five updates under a short exponential decay (x 0.5 after updates 2 and 4): every update
    equals the restatement's Adam step, at the restatement's learning-rate factor, on the
    trainer's previous parameters and its own gradients (moments carried in FP64). A first Adam
    step moves each parameter by about the learning rate, so a wrong factor shows as ~1e-3."""
    from oracle import train
    x, t = setup["x"], setup["t"]
    tr = nb.Trainer(ctx, setup["scene"], max_batch=4096, decay_start=2, decay_interval=2, decay_base=0.5)
    mom = {k: (np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64))
           for k, n in (("t", setup["tables"].size), ("m", setup["mlp"].size))}
    for step in range(1, 6):
        prev = {"t": tr.get(nb.TRAIN_TABLES).astype(np.float64), "m": tr.get(nb.TRAIN_MLP).astype(np.float64)}
        tr.step(x[step * 2000:step * 2000 + 2000], t[step * 2000:step * 2000 + 2000])
        grads = {"t": tr.get(nb.TRAIN_GRAD_TABLES).astype(np.float64), "m": tr.get(nb.TRAIN_GRAD_MLP).astype(np.float64)}
        got = {"t": tr.get(nb.TRAIN_TABLES), "m": tr.get(nb.TRAIN_MLP)}
        lr = 1e-2 * train.lr_factor(step, 2, 2, 0.5)
        for k in ("t", "m"):
            w, m, v, s = train.adam_step(prev[k], grads[k], mom[k][0], mom[k][1], mom[k][2], lr=lr,
                                         matrix=k == "m")
            mom[k] = (m, v, s)
            assert np.allclose(got[k], w, rtol=0, atol=1e-6), (step, k, np.abs(got[k] - w).max())
    tr.close()


def test_training_learns_visibility(nb, orc, ctx, setup):
    from nbvh_amd import hip
    x, t = setup["x"], setup["t"]
    rng = np.random.default_rng(1)
    held, pool = np.arange(20000), np.arange(20000, len(x))
    assert 0.1 < t[:, 0].mean() < 0.9

    def accuracy():
        out, _ = setup["mod"].network_inference(x[held], precision=nb.PREC_FP32)
        return float(np.mean((out[:, 0] > 0.5) == (t[held, 0] > 0.5)))

    acc0 = accuracy()
    scene = setup["scene"]
    tr = nb.Trainer(ctx, scene, max_batch=8192)
    dx, dt = hip.Buffer.from_array(x[pool]), hip.Buffer.from_array(t[pool])
    losses = []
    for step in range(250):
        off = int(rng.integers(0, len(pool) - 8192))
        losses.append(tr.step_device(dx.ptr + off * 36, dt.ptr + off * 32, 8192))
    tr.commit()
    acc1 = accuracy()
    prior = max(t[held, 0].mean(), 1 - t[held, 0].mean())
    # Measured: held-out accuracy 0.62 (the class prior) -> 0.86 after 250 steps, loss / 10.
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    assert acc1 > prior + 0.15 and acc1 > acc0 + 0.15, (acc0, acc1, prior)
    # The committed network is the trainer's, in every precision.
    cfg, tables, mlp = scene.network()
    assert np.array_equal(mlp, tr.get(nb.TRAIN_MLP)) and np.array_equal(tables, tr.get(nb.TRAIN_TABLES))
    out16, _ = setup["mod"].network_inference(x[held], precision=nb.PREC_FP16)
    assert np.mean((out16[:, 0] > 0.5) == (t[held, 0] > 0.5)) > acc1 - 0.01
    tr.close()
    # Put the module's seeded network back for the tests that compare against setup's copies.
    scene.load_network_model(setup["cfg"], setup["tables"], setup["mlp"])


def test_trainer_refuses_a_replaced_network(nb, orc, ctx, setup):
    """A trainer's buffers are sized for the network of its creation: after the scene's network
    is replaced (here by a smaller table size), step and commit fail with NBVH_ERR_STATE before
    any launch or copy, and a new trainer works on the new network."""
    v, tri, _, _, _ = bench_scene(nb)
    scene = nb.Scene(ctx, v, tri)
    scene.build_nbvh(64)
    scene.init_network(nb.network_config(), seed=2, table_range=1e-2)
    x, t = setup["x"][:512], setup["t"][:512]
    tr = nb.Trainer(ctx, scene, max_batch=1024)
    tr.step(x, t)
    scene.init_network(nb.network_config(log2_hashmap_size=14), seed=3, table_range=1e-2)
    for call in (lambda: tr.step(x, t), tr.commit):
        with pytest.raises(nb.NbvhError) as e:
            call()
        assert e.value.status == nb._ffi.ERR_STATE
    tr.close()
    tr2 = nb.Trainer(ctx, scene, max_batch=1024)
    assert np.isfinite(tr2.step(x, t))
    tr2.commit()
    tr2.close()
    scene.close()


def test_trainer_refuses_batches_beyond_32bit_term_indices(nb, orc, ctx, setup):
    """The table gradient indexes a step's corner terms with 32 bits: a max batch whose
    n_points * n_levels * 8 terms per sample reach 2^32 (2^24 samples at 3 points x 16 levels:
    6.4e9 terms) is refused at creation with NBVH_ERR_INVALID_ARGUMENT, before any allocation."""
    with pytest.raises(nb.NbvhError) as e:
        nb.Trainer(ctx, setup["scene"], max_batch=1 << 24)
    assert e.value.status == nb._ffi.ERR_INVALID_ARGUMENT
