"""CPU restatement of the reference's tree-cut learning (test infrastructure only; never used by
the product): the split schedule, node losses and heuristics, split candidates, splits and LoD
assignment of the training loop, src/cuda/neural/neural_bvh_renderer.cu:141-297.

  SplitSchedule     BvhSplitScheduler, include/cuda/neural/utils/bvh_split_scheduler.h:13-144
  initial_cut       traverse_and_build_neural_bvh, src/cuda/neural/neural_bvh_renderer_accel.cpp:25-88
  TreeCut.update    update_neural_tree_cut_loss / update_neural_tree_cut_decisions,
                    src/cuda/neural/neural_bvh_renderer_kernels.cu:202-288, and the loop's split and
                    LoD steps (neural_bvh_renderer.cu:160-276)
  TreeCut.split     split_neural_bvh_at, neural_bvh_renderer_accel.cpp:256-392
  TreeCut.assign_lod assign_lod_to_current_neural_leaf_nodes, :395-440

Conventions where the reference leaves the result to a race: the per-step node loss is the FP64
sum of the samples' losses, added once to the FP32 node loss; among leaves of equal heuristic the
lowest node index is the split candidate.
"""
import numpy as np

NO_CHILD = 0xFFFFFFFF
LEAF_FINAL = 1
NUM_LODS = 7


def leaf_lod(lod):
    return 2 + lod


def w32(x):
    """x wrapped to int32 (two's complement), as the reference's compiled int arithmetic."""
    return ((int(x) + (1 << 31)) % (1 << 32)) - (1 << 31)


def mul_trunc(a, f):
    """int *= float as x86's cvttss2si: truncation toward zero, INT_MIN out of range."""
    p = np.float32(a) * np.float32(f)
    return int(p) if -2147483648.0 <= p < 2147483648.0 else -(1 << 31)


def c_div(a, b):
    """C integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


class SplitSchedule:
    """BvhSplitScheduler with int truncation of the float scalings (int *= float) and its int
    arithmetic wrapping in two's complement (its node counts overflow within shipped schedules)."""

    def __init__(self, splitting_interval=50, last_split_iteration=2000, start_split_num=1, split_scaling=2.0,
                 split_interval_scaling=2.0, lod_decision=0):
        self.interval, self.last, self.start = splitting_interval, last_split_iteration, start_split_num
        self.scaling, self.iscaling = np.float32(split_scaling), np.float32(split_interval_scaling)
        self.lod_decision = lod_decision
        self.stats_total = 0
        self.reset()

    def reset(self):
        self.num, self.cur_interval, self.total = self.start, self.interval, 0

    def splits(self, it):
        if it == 0:
            self.reset()
        if it > self.last:
            return 0
        n = 0
        if it % self.interval == 0:
            n = self.num
            self.total = w32(self.total + n)
        if it % self.cur_interval == 0:
            self.num = mul_trunc(self.num, self.scaling)
            self.cur_interval = mul_trunc(self.cur_interval, self.iscaling)
        return n

    def update_stats_and_reset(self, start_nodes, num_iterations):
        self.reset()
        self.stats_total = 0
        nodes = start_nodes
        for i in range(num_iterations):
            leaves = c_div(w32(nodes + 1), 2)
            s = self.splits(i)
            if s > 0:
                s = min(s, leaves)
                self.stats_total = w32(self.stats_total + s)
                nodes = w32(nodes + w32(2 * s))
        self.reset()

    def should_assign_new_lod(self, it, current_lod, num_lods_learned, max_lods=NUM_LODS):
        if self.lod_decision == 0:
            assign = c_div(w32((max_lods - current_lod) * self.last), num_lods_learned) <= it
        else:
            assign = c_div(w32((max_lods - current_lod) * self.stats_total), num_lods_learned) <= self.total
        return assign or it >= self.last


def heuristic(loss, iter_alive, samples):
    """NeuralTreeCutData::compute_loss_heuristic in float32."""
    if samples == 0:
        return np.float32(-np.finfo(np.float32).max)
    f = np.float32
    return f(f(loss) / f(f(iter_alive) * f(np.sqrt(f(f(samples) / f(iter_alive))))))


def initial_cut(bvh, max_depth, first_lod=NUM_LODS - 1):
    """The triangle BVH copied down to max_depth, children numbered in pairs as they are created."""
    nodes, node_map, depth = [None], [0], [0]

    def rec(bi, ni, d):
        b = bvh[bi]
        node_map[ni], depth[ni] = bi, d
        leaf = b["prim_count"] > 0
        if leaf or d == max_depth:
            nodes[ni] = (b["aabb_min"].copy(), LEAF_FINAL if leaf else leaf_lod(first_lod), b["aabb_max"].copy(),
                         NO_CHILD)
            return
        left = len(nodes)
        nodes.extend([None, None])
        node_map.extend([0, 0])
        depth.extend([0, 0])
        nodes[ni] = (b["aabb_min"].copy(), 0, b["aabb_max"].copy(), left)
        rec(int(b["left_or_prim_offset"]), left, d + 1)
        rec(int(b["left_or_prim_offset"]) + 1, left + 1, d + 1)

    rec(0, 0, 0)
    out = np.zeros(len(nodes), bvh.dtype)
    for i, (mn, pc, mx, left) in enumerate(nodes):
        out[i]["aabb_min"], out[i]["prim_count"], out[i]["aabb_max"], out[i]["left_or_prim_offset"] = mn, pc, mx, left
    return out, node_map, depth


class TreeCut:
    """The learner's state and one step of its tree-cut update."""

    def __init__(self, bvh, min_depth=0, num_lods_learned=1, tree_cut_learning=True, max_training_steps=10000,
                 **schedule):
        self.bvh = bvh
        self.nodes, self.node_map, depth = initial_cut(bvh, min_depth)
        n = len(self.nodes)
        self.loss = np.zeros(n, np.float32)
        self.iter_alive = np.ones(n, np.uint32)
        self.depth = np.array(depth, np.uint32)
        self.samples = np.zeros(n, np.uint32)
        self.sched = SplitSchedule(**schedule)
        self.num_lods_learned = num_lods_learned
        self.tree_cut = bool(tree_cut_learning)
        self.current_lod = NUM_LODS - 1
        self.all_assigned = False
        self.current_num_splits = max(1, self.sched.start)
        self.sched.update_stats_and_reset(n, max_training_steps)
        self.step = 0
        self.max_node_loss = np.float32(0)
        self.split_idx = []
        self.total_splits = 0

    def tree_cut_data(self):
        return self.loss, self.iter_alive, self.depth, self.samples

    def update(self, node_idx, sample_loss):
        should_split = False
        if self.tree_cut:
            n = self.sched.splits(self.step)
            if n != self.current_num_splits:
                self.current_num_splits = max(1, n)
            should_split = n > 0
        # update_neural_tree_cut_loss: every sample's loss to its node, 8 samples per slot
        sums = np.zeros(len(self.nodes), np.float64)
        counts = np.zeros(len(self.nodes), np.uint64)
        np.add.at(sums, np.asarray(node_idx, np.int64), np.asarray(sample_loss, np.float32).astype(np.float64))
        np.add.at(counts, np.asarray(node_idx, np.int64), 8)
        touched = counts > 0
        self.loss[touched] = (self.loss[touched] + sums[touched].astype(np.float32)).astype(np.float32)
        self.samples[touched] += counts[touched].astype(np.uint32)
        # update_neural_tree_cut_decisions: leaves age; splittable leaves' heuristics
        keys = np.zeros(len(self.nodes), np.float32)
        for i in range(len(self.nodes)):
            if self.nodes[i]["left_or_prim_offset"] != NO_CHILD:
                continue
            self.iter_alive[i] += 1
            if self.nodes[i]["prim_count"] == LEAF_FINAL:
                continue
            h = heuristic(self.loss[i], self.iter_alive[i], self.samples[i])
            keys[i] = h if h > 0 else np.float32(0)
        order = sorted(range(len(keys)), key=lambda i: (-float(keys[i]), i))
        k = self.current_num_splits
        self.split_idx = [0] * k
        found, prev = 0, None
        for i in order:
            if found == k or not keys[i] > 0:
                break
            if prev is not None and keys[i] == prev:
                continue
            self.split_idx[found] = i
            found += 1
            prev = keys[i]
        self.max_node_loss = keys[order[0]] if len(order) and keys[order[0]] > 0 else np.float32(0)
        if self.tree_cut:
            if should_split:
                self.split()
            if self.num_lods_learned > 1 and not self.all_assigned:
                if self.sched.should_assign_new_lod(self.step, self.current_lod, self.num_lods_learned):
                    self.assign_lod()
                if self.all_assigned:
                    self.current_num_splits = 1
        self.step += 1

    def split(self):
        leaves = (len(self.nodes) + 1) // 2
        unique = sorted(set(self.split_idx[:min(leaves, len(self.split_idx))]))
        new_nodes, new_loss, new_ia, new_depth, new_samples = [], [], [], [], []
        nodes = [tuple(n) for n in self.nodes]
        for s in unique:
            parent = self.bvh[self.node_map[s]]
            if parent["prim_count"] > 0 or nodes[s][3] != NO_CHILD:
                continue
            left = len(nodes)
            pl = int(parent["left_or_prim_offset"])
            self.node_map += [pl, pl + 1]
            mn, pc, mx, _ = nodes[s]
            if pc == leaf_lod(self.current_lod):
                pc = 0
            nodes[s] = (mn, pc, mx, left)
            d = int(self.depth[s]) + 1
            for c in range(2):
                tb = self.bvh[pl + c]
                nodes.append((tb["aabb_min"].copy(), LEAF_FINAL if tb["prim_count"] > 0 else leaf_lod(self.current_lod),
                              tb["aabb_max"].copy(), NO_CHILD))
                new_loss.append(0.0)
                new_ia.append(1)
                new_depth.append(d)
                new_samples.append(0)
            self.total_splits += 1
        out = np.zeros(len(nodes), self.nodes.dtype)
        for i, (mn, pc, mx, left) in enumerate(nodes):
            out[i]["aabb_min"], out[i]["prim_count"], out[i]["aabb_max"], out[i]["left_or_prim_offset"] = mn, pc, mx, left
        self.nodes = out
        self.loss = np.concatenate([self.loss, np.array(new_loss, np.float32)])
        self.iter_alive = np.concatenate([self.iter_alive, np.array(new_ia, np.uint32)])
        self.depth = np.concatenate([self.depth, np.array(new_depth, np.uint32)])
        self.samples = np.concatenate([self.samples, np.array(new_samples, np.uint32)])

    def assign_lod(self):
        lod = self.current_lod
        if lod != NUM_LODS - 1:
            for i in range(len(self.nodes)):
                if self.nodes[i]["left_or_prim_offset"] == NO_CHILD:
                    pc = self.nodes[i]["prim_count"]
                    self.nodes[i]["prim_count"] = LEAF_FINAL if pc == LEAF_FINAL else leaf_lod(lod)
        self.all_assigned = lod == 0
        if not self.all_assigned:
            self.current_lod -= 1
        else:
            self.tree_cut = False
