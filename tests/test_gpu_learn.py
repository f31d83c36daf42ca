"""Tree-cut learning on the GPU (nbvh_learner_*) against its restatements.

  - the sampler kernel's batches (encoded inputs, targets, sampled nodes), step after step, are
    bit-identical to oracle.learn_sample given the same N-BVH, tree-cut data and previous slots;
  - the tree-cut update (node losses, ageing, heuristics, candidates, splits, LoD assignment) from
    the same per-slot losses matches oracle.learn.TreeCut: the N-BVH and every integer of the
    tree-cut data exactly, node losses within 1e-6 relative (FP64 step sums on both sides);
  - an end-to-end run (sample, train, split) grows the N-BVH from its root and trains a network
    whose FP16 query flips at most 0.5 % of the FP64 oracle's visibility decisions.
"""
import numpy as np
import pytest

from conftest import camera_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(nb):
    c = nb.Context()
    yield c
    c.close()


def make_scene(nb, ctx, subdiv=5, seed=3):
    v, t = nb.icosphere(subdiv, seed=seed, amplitude=0.1)
    scene = nb.Scene(ctx, v, t)
    scene.init_network(nb.network_config(log2_hashmap_size=16), seed=5, table_range=1e-4)
    return v, t, scene


def oracle_sample(orc, learner, scene, v, t, step, tc, prev, max_loss, n_sub_batch=1):
    from nbvh_amd import _ffi
    nodes, prims = scene.bvh()
    bmin, bmax = scene.bbox()
    st = learner.state()
    return orc.learn_sample(learner.get(_ffi.LEARN_NBVH), tc, v, t, nodes, prims, bmin, bmax, step=step,
                            batch_size=learner.config.batch_size, prev=prev, n_sub_batch=n_sub_batch,
                            current_lod=st["current_learned_lod"], survival=True,
                            current_max_loss=max_loss if step > 0 else 0.0)


@pytest.mark.parametrize("n_sub_batch_log2,lod_decision", [(0, 0), (2, 0), (0, 1)])
def test_sampler_and_tree_cut_match_the_restatements(nb, orc, ctx, n_sub_batch_log2, lod_decision):
    """The schedule grows its split count 1.5x per step, so its expected node count leaves the int
    range within the 60 planned steps: the LoD decisions by equal splits (lod_decision 1) then rest
    on the wrapped totals, on both sides."""
    from nbvh_amd import _ffi
    from oracle import learn as OL
    v, t, scene = make_scene(nb, ctx)
    batch = 1 << 15
    sched = dict(splitting_interval=1, last_split_iteration=50, start_split_num=3, split_scaling=1.5,
                 split_interval_scaling=1.0, lod_decision=lod_decision)
    lr = nb.Learner(ctx, scene, batch_size=batch, neural_bvh_min_depth=4, log2_bvh_max_traversals=n_sub_batch_log2,
                    num_lods_learned=2, max_training_steps=60, **sched)
    nodes, _ = scene.bvh()
    ref = OL.TreeCut(nodes, min_depth=4, num_lods_learned=2, max_training_steps=60, **sched)
    assert np.array_equal(lr.get(_ffi.LEARN_NBVH).view(np.uint32), ref.nodes.view(np.uint32))
    assert np.array_equal(scene.nbvh().view(np.uint32), ref.nodes.view(np.uint32))
    prev = None
    rng = np.random.default_rng(7)
    for step in range(8):
        print(f"[learner] step {step}: {len(ref.nodes)} nodes", flush=True)
        tc = np.zeros(len(ref.nodes), orc.TREE_CUT_DTYPE)
        tc["loss"], tc["iter_alive"], tc["depth"], tc["samples"] = ref.tree_cut_data()
        exp = oracle_sample(orc, lr, scene, v, t, step, tc, prev, float(ref.max_node_loss), 1 << n_sub_batch_log2)
        lr.sample(step)
        got_in, got_tg, got_nd = lr.get(_ffi.LEARN_INPUTS), lr.get(_ffi.LEARN_TARGETS), lr.get(_ffi.LEARN_NODES)
        assert np.array_equal(got_nd, exp["node_idx"]), step
        assert np.array_equal(got_in.view(np.uint32), exp["inputs"].view(np.uint32)), step
        assert np.array_equal(got_tg.view(np.uint32), exp["targets"].view(np.uint32)), step
        if step > 0:
            assert 0.0 < exp["written"].mean() < 1.0  # survival drops some samples
        prev = (exp["inputs"], exp["targets"], exp["node_idx"])
        # per-slot losses: a seeded function of the slot's node (so some leaves stand out)
        loss = (rng.uniform(0.5, 1.5, batch) * (1.0 + (got_nd % 5))).astype(np.float32) * np.float32(1e-6)
        lr.update(loss)
        ref.update(exp["node_idx"], loss)
        st = lr.state()
        assert st["n_nodes"] == len(ref.nodes) and st["current_learned_lod"] == ref.current_lod
        assert np.array_equal(lr.get(_ffi.LEARN_SPLITS), np.array(ref.split_idx, np.uint32)), step
        assert np.array_equal(lr.get(_ffi.LEARN_NBVH).view(np.uint32), ref.nodes.view(np.uint32)), step
        g = lr.get(_ffi.LEARN_TREE_CUT)
        for f in ("iter_alive", "depth", "samples"):
            assert np.array_equal(g[f], getattr(ref, {"iter_alive": "iter_alive", "depth": "depth",
                                                      "samples": "samples"}[f])), (step, f)
        assert np.allclose(g["loss"], ref.loss, rtol=1e-6, atol=0)
        assert st["max_node_loss"] == pytest.approx(float(ref.max_node_loss), rel=1e-6)
        assert st["sample_failures"] == 0
    assert len(ref.nodes) > 31  # splits happened
    assert np.array_equal(scene.nbvh().view(np.uint32), ref.nodes.view(np.uint32))
    lr.close()
    scene.close()


def test_learning_grows_the_nbvh_and_trains(nb, orc, ctx):
    v, t, scene = make_scene(nb, ctx, subdiv=5, seed=11)
    lr = nb.Learner(ctx, scene, batch_size=1 << 16, max_training_steps=400, splitting_interval=20,
                    last_split_iteration=300)
    tr = nb.Trainer(ctx, scene, max_batch=1 << 16)
    assert lr.state()["n_nodes"] == 1
    losses = [lr.step(tr) for _ in range(400)]
    st = lr.state()
    print(f"[learner] {st}", flush=True)
    assert st["sample_failures"] == 0 and st["step"] == 400
    assert st["n_nodes"] > 50 and st["total_splits"] == (st["n_nodes"] - 1) // 2
    assert np.mean(losses[-20:]) < 0.5 * np.mean(losses[:5]), (losses[:5], losses[-20:])
    tr.commit()
    # the trained network on the learned N-BVH: FP16 visibility against the FP64 oracle
    mod = nb.NeuralBVHModule(ctx, scene)
    _, rays = camera_rays(nb, orc, 160, 120)
    cfg, tables, mlp = scene.network()
    bmin, bmax = scene.bbox()
    net = orc.Net(cfg, tables, mlp, bmin, bmax)
    e64 = net.neural_query(scene.nbvh(), rays, mode=orc.FP64)
    g16 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP16)
    g32 = mod.neural_bvh_traversal(rays, precision=nb.PREC_FP32)
    assert np.array_equal(g32.view(np.uint32), net.neural_query(scene.nbvh(), rays, mode=orc.FP32).view(np.uint32))
    flips = float(np.mean(nb.hit_kind(g16) != nb.hit_kind(e64)))
    print(f"[learner] trained network: FP16 visibility flips {flips:.4%}, hit fraction "
          f"{(nb.hit_kind(e64) != 0).mean():.3f}", flush=True)
    assert (nb.hit_kind(e64) != 0).mean() > 0.05
    assert flips <= 0.005
    tr.close()
    lr.close()
    scene.close()
