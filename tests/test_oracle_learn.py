"""The restatements of the training-data sampler and of tree-cut learning, pinned on the CPU.

Sampler (oracle.learn_sample): its float sin/cos against FP64; directions unit length with the
moments of the uniform sphere; origins inside the scaled scene box; every sampled segment is the
slab of its node dilated by its sampled dilation, accepted by the learning epsilons and a leaf at
its LoD; the targets equal the independent segment_labels; the inputs the closed-form encoding
(no swap of a negative entry); survival keeps the expected fraction; a segment without a hit is
followed by the next leaf of the same ray.
Tree cut (oracle/learn.py): the split schedule worked by hand; the heuristic's closed form; the
initial cut against the triangle BVH; candidate choice, splits and LoD assignment on small cases
worked by hand.
"""
import numpy as np
import pytest

from oracle import learn as L


@pytest.fixture(scope="module")
def scene(nb, small_scene):
    v, t, nodes, prims, nbn = small_scene
    return dict(v=v, t=t, nodes=nodes, prims=prims, nbvh=nbn, bmin=nodes[0]["aabb_min"], bmax=nodes[0]["aabb_max"])


def fresh_tree_cut(n):
    import oracle
    tc = np.zeros(n, oracle.TREE_CUT_DTYPE)
    tc["iter_alive"] = 1
    return tc


def test_sincos_against_fp64(orc):
    x = np.concatenate([np.linspace(0, 2 * np.pi, 200001), [0.0, np.pi / 2, np.pi, 3 * np.pi / 2]]).astype(np.float32)
    s, c = orc.sincosf(x)
    xd = x.astype(np.float64)
    for got, ref in ((s, np.sin(xd)), (c, np.cos(xd))):
        err = np.abs(got.astype(np.float64) - ref)
        assert err.max() < 1.2e-7
        big = np.abs(ref) > 1e-2
        ulp = np.spacing(np.abs(ref[big]).astype(np.float32)).astype(np.float64)
        assert (err[big] / ulp).max() <= 2.0


def test_sampler_segments(nb, orc, scene):
    n = 8192
    r = orc.learn_sample(scene["nbvh"], fresh_tree_cut(len(scene["nbvh"])), scene["v"], scene["t"], scene["nodes"],
                         scene["prims"], scene["bmin"], scene["bmax"], step=3, batch_size=n, survival=False)
    assert r["written"].all()
    d = r["rays"]["d"].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    assert np.abs(d.mean(axis=0)).max() < 0.03 and np.abs((d ** 2).mean(axis=0) - 1 / 3).max() < 0.02
    ext = scene["bmax"] - scene["bmin"]
    o = r["rays"]["o"]
    assert (o >= scene["bmin"] - 0.25 * ext - 1e-6).all() and (o <= scene["bmax"] + 0.25 * ext + 1e-6).all()
    assert (r["dil"] >= 1e-5).all() and (r["dil"] <= 1.1e-4 + 1e-9).all() and (r["lod"] == 6).all()
    nodes = scene["nbvh"]
    for i in range(0, n, 7):
        node = nodes[r["node_idx"][i]]
        assert node["left_or_prim_offset"] == 0xFFFFFFFF or 0 < node["prim_count"] <= 2 + r["lod"][i]
        dmin, dmax = orc.dilate(node, float(r["dil"][i]))
        hit, tn, tf = orc.slab(dmin, dmax, r["rays"][i])
        mint, maxt = r["seg"][i]
        assert hit and (np.float32(tn), np.float32(tf)) == (mint, maxt)
        assert maxt > 1e-3 if mint < 0 else mint > 1e-3
    # targets: the independent label restatement on the same segments
    lab = orc.segment_labels(scene["v"], scene["t"], scene["nodes"], scene["prims"], r["rays"], r["seg"][:, 0],
                             r["seg"][:, 1])
    assert np.array_equal(lab.view(np.uint32), r["targets"].view(np.uint32))
    assert 0.05 < r["targets"][:, 0].mean() < 0.95
    # inputs: entry-to-exit points at t = 0, 1/2, 1 of the encoding box (no swap)
    f = np.float32
    c = (scene["bmax"] + scene["bmin"]) * f(0.5)
    e = (scene["bmax"] - scene["bmin"]) * ((f(1) + f(1e-4)) + f(1e-5))
    mint, maxt = r["seg"][:, :1], r["seg"][:, 1:]
    eo = (r["rays"]["o"] + mint * r["rays"]["d"] - c) / e + f(0.5)
    ei = (r["rays"]["o"] + maxt * r["rays"]["d"] - c) / e + f(0.5)
    for k, t in enumerate((f(0), f(0.5), f(1))):
        assert np.array_equal(r["inputs"][:, 3 * k:3 * k + 3], (f(1) - t) * eo + t * ei)


def test_sampler_continues_rays_without_hits(orc, scene):
    n = 4096
    r = orc.learn_sample(scene["nbvh"], fresh_tree_cut(len(scene["nbvh"])), scene["v"], scene["t"], scene["nodes"],
                         scene["prims"], scene["bmin"], scene["bmax"], step=1, batch_size=n, n_sub_batch=4,
                         survival=False)
    rays = r["rays"].reshape(-1, 4)
    vis = r["targets"][:, 0].reshape(-1, 4)
    node = r["node_idx"].reshape(-1, 4)
    same = miss = 0
    for j in range(len(rays)):
        for k in range(3):
            if vis[j, k] == 0:
                miss += 1
                # the next sample continues the same ray to its next leaf, or restarts
                if (np.array_equal(rays[j, k + 1]["o"], rays[j, k]["o"]) and
                        np.array_equal(rays[j, k + 1]["d"], rays[j, k]["d"])):
                    same += 1
                    assert node[j, k + 1] != node[j, k]  # the traversal went on to another leaf
    assert miss > 100 and same > 0.3 * miss


def test_sampler_survival_fraction(orc, scene):
    nbn = scene["nbvh"]
    tc = fresh_tree_cut(len(nbn))
    tc["samples"], tc["iter_alive"], tc["loss"] = 8, 2, 1.0  # heuristic 1 / (2 sqrt 4) = 0.25
    n = 8192
    kw = dict(step=2, batch_size=n, survival=True)
    args = (nbn, tc, scene["v"], scene["t"], scene["nodes"], scene["prims"], scene["bmin"], scene["bmax"])
    half = orc.learn_sample(*args, current_max_loss=0.5, **kw)
    assert abs(half["written"].mean() - 0.5) < 0.03
    # above the previous maximum: always kept; far below: kept with min_rr_survival_prob
    assert orc.learn_sample(*args, current_max_loss=0.2, **kw)["written"].all()
    low = orc.learn_sample(*args, current_max_loss=1e6, **kw)["written"].mean()
    assert low < 0.02
    # a dropped slot keeps its previous inputs and targets, its node index is the new sample's
    prev = (np.full((n, 9), 7.0, np.float32), np.full((n, 8), 3.0, np.float32), np.zeros(n, np.uint32))
    r = orc.learn_sample(*args, current_max_loss=0.5, prev=prev, **kw)
    dropped = r["written"] == 0
    assert (r["inputs"][dropped] == 7.0).all() and (r["targets"][dropped] == 3.0).all()
    assert (r["node_idx"][dropped] == half["node_idx"][dropped]).all()


def test_split_schedule_by_hand():
    s = L.SplitSchedule(splitting_interval=50, last_split_iteration=400, start_split_num=1, split_scaling=2.0,
                        split_interval_scaling=2.0)
    got = {it: s.splits(it) for it in range(0, 460)}
    nonzero = {it: n for it, n in got.items() if n}
    # splits every 50 iterations; the count doubles at 0, 100, 200, 400 (the interval doubles too)
    assert nonzero == {0: 1, 50: 2, 100: 2, 150: 4, 200: 4, 250: 8, 300: 8, 350: 8, 400: 8}
    t = L.SplitSchedule(splitting_interval=40, last_split_iteration=3000, start_split_num=1, split_scaling=2.5,
                        split_interval_scaling=2.0)
    seq = [t.splits(it) for it in range(0, 200)]
    # int truncation: 1 -> 2 (2.5) -> 5 -> 12 (12.5)
    assert [seq[i] for i in (0, 40, 80, 120, 160)] == [1, 2, 2, 5, 5]
    s.update_stats_and_reset(1, 460)
    # splits are capped by the leaves: 1 node -> 1 leaf at it 0, then (1+2*1+1)//2 = 2 at it 50 ...
    assert s.stats_total == 1 + 2 + 2 + 4 + 4 + 8 + 8 + 8 + 8
    assert s.should_assign_new_lod(399, 6, 1) is False and s.should_assign_new_lod(400, 6, 1) is True
    assert s.should_assign_new_lod(200, 6, 2) is True and s.should_assign_new_lod(199, 6, 2) is False


def test_heuristic_closed_form():
    assert L.heuristic(1.0, 2, 8) == np.float32(0.25)
    assert L.heuristic(3.0, 4, 64) == np.float32(3.0 / (4 * 4))
    assert L.heuristic(5.0, 1, 0) == -np.finfo(np.float32).max


def test_initial_cut_and_splits(nb, scene):
    bvh = scene["nodes"]
    nodes, node_map, depth = L.initial_cut(bvh, 2)
    # children in pairs as created, depth first: 0 -> (1, 2), 1 -> (3, 4), 2 -> (5, 6)
    assert len(nodes) == 7 and depth == [0, 1, 1, 2, 2, 2, 2]
    assert list(nodes["left_or_prim_offset"][:3]) == [1, 3, 5]
    for i in range(7):
        assert np.array_equal(nodes[i]["aabb_min"], bvh[node_map[i]]["aabb_min"])
    assert (nodes["prim_count"][3:] == L.leaf_lod(6)).all()
    # a learner on the root alone: step 0 splits the only leaf (positive loss)
    tc = L.TreeCut(bvh, min_depth=0, max_training_steps=100)
    assert len(tc.nodes) == 1 and tc.nodes[0]["prim_count"] == L.leaf_lod(6)
    tc.update(np.zeros(16, np.uint32), np.full(16, 0.5, np.float32))
    assert tc.loss[0] == np.float32(8.0) and tc.samples[0] == 128 and tc.iter_alive[0] == 2
    assert tc.split_idx == [0] and len(tc.nodes) == 3 and tc.nodes[0]["prim_count"] == 0
    assert list(tc.depth) == [0, 1, 1] and list(tc.iter_alive[1:]) == [1, 1]
    assert tc.max_node_loss == L.heuristic(8.0, 2, 128)


def test_candidates_ties_and_lod(nb, scene):
    bvh = scene["nodes"]
    tc = L.TreeCut(bvh, min_depth=2, num_lods_learned=2, splitting_interval=1, last_split_iteration=10,
                   start_split_num=2, split_scaling=1.0, split_interval_scaling=1.0, max_training_steps=20)
    # leaves 3..6; equal losses on 3 and 5 and a larger one on 6: candidates 6, then 3 (ties -> lowest)
    tc.update(np.array([3, 5, 6, 6], np.uint32), np.array([1.0, 1.0, 1.0, 1.0], np.float32))
    assert tc.split_idx == [6, 3]
    # splits run in ascending node order: node 3 gets children 7, 8 and node 6 gets 9, 10
    assert tc.nodes[3]["left_or_prim_offset"] == 7 and tc.nodes[6]["left_or_prim_offset"] == 9
    # EqualIterations with 2 LoDs: (7 - 6) * 10 // 2 = 5 -> LoD 6 assigned at step 5, then LoD 5 learned
    for _ in range(5):
        tc.update(np.array([4], np.uint32), np.array([0.1], np.float32))
    assert tc.current_lod == 5 and not tc.all_assigned
    leaves = tc.nodes["left_or_prim_offset"] == L.NO_CHILD
    assert set(tc.nodes["prim_count"][leaves]) <= {L.LEAF_FINAL, L.leaf_lod(6)}
    n_before = len(tc.nodes)
    for _ in range(3):  # losses on every current leaf: splits go on, now at LoD 5
        leaf_ids = np.nonzero(tc.nodes["left_or_prim_offset"] == L.NO_CHILD)[0].astype(np.uint32)
        tc.update(leaf_ids, np.linspace(1.0, 2.0, len(leaf_ids)).astype(np.float32))
    assert len(tc.nodes) > n_before
    new = tc.nodes[n_before:]
    assert set(new["prim_count"][new["left_or_prim_offset"] == L.NO_CHILD]) <= {L.LEAF_FINAL, L.leaf_lod(5)}


def test_split_schedule_int_arithmetic():
    """The scheduler's int arithmetic as compiled code performs it: split counts doubling every
    iteration reach 2^30, then int *= float leaves the int range (INT_MIN, as x86's truncating
    conversion gives), and the running total wraps in two's complement."""
    from oracle.learn import SplitSchedule
    s = SplitSchedule(splitting_interval=1, last_split_iteration=40, start_split_num=1, split_scaling=2.0,
                      split_interval_scaling=1.0)
    got = [s.splits(i) for i in range(36)]
    assert got[:31] == [1 << i for i in range(31)]
    assert got[31:] == [-(1 << 31)] * 5
    assert s.total == ((sum(got) + (1 << 31)) % (1 << 32)) - (1 << 31)
    # Expected nodes from 1: split i adds 2^i (the leaf cap equals it), so after i = 29 the count is
    # 2^31 - 1; at i = 30, (nodes + 1) / 2 wraps to INT_MIN / 2, the cap turns the 2^30 splits into
    # -2^30, and the total 2^30 - 1 becomes -1; later counts are negative (no splits).
    s.update_stats_and_reset(1, 36)
    assert s.stats_total == -1
