// learn.cpp -- nbvh_learner: tree-cut learning of the N-BVH around the trainer (the reference's
// training loop, src/cuda/neural/neural_bvh_renderer.cu:141-297). The sampler and the node-loss
// kernels are in kernels/learn.hip; the split schedule, the splits and the LoD assignment run on
// the host, as in the reference, once per step after one synchronisation.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <vector>

#include "../kernels/launch.h"
#include <climits>

#include "context.h"

using namespace nbvh;

namespace {

// The scheduler's int arithmetic as the reference's compiled code performs it: its expected node
// count and split totals overflow int within its own shipped schedules (e.g. interval 40, last
// 3000, scaling 2.5), which wraps in two's complement; int *= float truncates like x86's
// cvttss2si, INT_MIN when the product is out of range.
int wrap_add(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
int wrap_mul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
int mul_trunc(int a, float f)
{
    const float p = (float)a * f;
    return (p >= -2147483648.f && p < 2147483648.f) ? (int)p : INT_MIN;
}

// BvhSplitScheduler (include/cuda/neural/utils/bvh_split_scheduler.h:13-144). The products with
// the float scalings truncate to int, as the reference's int *= float does.
struct SplitScheduler {
    int interval = 50, last_iteration = 2000, start_num = 1;
    float scaling = 2.f, interval_scaling = 2.f;
    uint32_t lod_decision = 0;  // 0 EqualIterations, 1 EqualSplits
    int cur_num = 1, cur_interval = 50, cur_total = 0;
    int stats_total_splits = 0;

    void reset()
    {
        cur_num      = start_num;
        cur_interval = interval;
        cur_total    = 0;
    }
    int splits(int it)  // get_bvh_splits, :67-90
    {
        if (it == 0) reset();
        if (it > last_iteration) return 0;
        int n = 0;
        if (it % interval == 0) {
            n         = cur_num;
            cur_total = wrap_add(cur_total, n);
        }
        if (it % cur_interval == 0) {
            cur_num      = mul_trunc(cur_num, scaling);
            cur_interval = mul_trunc(cur_interval, interval_scaling);
        }
        return n;
    }
    void update_stats_and_reset(int start_nodes, int num_iterations)  // :39-65
    {
        reset();
        stats_total_splits = 0;
        int expected_nodes = start_nodes;
        for (int i = 0; i < num_iterations; i++) {
            const int leaves = wrap_add(expected_nodes, 1) / 2;
            int s = splits(i);
            if (s > 0) {
                if (s > leaves) s = leaves;
                stats_total_splits = wrap_add(stats_total_splits, s);
                expected_nodes     = wrap_add(expected_nodes, wrap_mul(2, s));
            }
        }
        reset();
    }
    bool should_assign_new_lod(int it, int current_lod, int num_lods_learned, int max_lods) const  // :92-118
    {
        bool assign = false;
        if (lod_decision == 0)
            assign = wrap_mul(max_lods - current_lod, last_iteration) / num_lods_learned <= it;
        else
            assign = wrap_mul(max_lods - current_lod, stats_total_splits) / num_lods_learned <= cur_total;
        return assign || it >= last_iteration;
    }
};

bool is_tri_leaf(const nbvh_node& n) { return n.prim_count > 0u; }

}  // namespace

struct nbvh_learner {
    nbvh_ctx* ctx     = nullptr;
    nbvh_scene* scene = nullptr;
    nbvh_learner_config cfg{};
    std::vector<nbvh_node> nodes;  // the N-BVH as learned (boxes of the triangle BVH nodes)
    std::vector<uint32_t> node_map;  // N-BVH node -> triangle BVH node
    std::vector<uint32_t> depth;     // N-BVH node depth (tree-cut data)
    uint32_t capacity = 0;           // device node arrays' capacity
    DevBuf d_nodes, d_tc, d_sums, d_counts, d_keys, d_vals, d_keys_sorted, d_vals_sorted, d_sort_tmp;
    size_t sort_tmp_bytes = 0;
    DevBuf d_inputs, d_targets, d_node_idx, d_sample_loss, d_failures;
    SplitScheduler sched;
    uint32_t step = 0, current_lod = NBVH_NUM_LODS - 1, current_num_splits = 1, total_splits = 0;
    bool all_assigned = false, tree_cut = true;
    float max_node_loss = 0.f, last_loss = 0.f;
    std::vector<uint32_t> split_idx;
};

namespace {

uint32_t n_inputs(const nbvh_learner* l) { return 3u * l->scene->cfg.n_points; }

// traverse_and_build_neural_bvh (src/cuda/neural/neural_bvh_renderer_accel.cpp:25-88): the
// triangle BVH copied down to max_depth, children numbered in pairs as they are created
// (depth first); cut nodes become LEAF_LOD(first_lod) leaves, triangle-BVH leaves LEAF_FINAL.
void build_cut(nbvh_learner* l, uint32_t bi, uint32_t ni, uint32_t d, uint32_t max_depth, uint32_t first_lod)
{
    const std::vector<nbvh_node>& bvh = l->scene->bvh;
    const nbvh_node& b = bvh[bi];
    l->node_map[ni] = bi;
    l->depth[ni]    = d;
    nbvh_node& n = l->nodes[ni];
    std::memcpy(n.aabb_min, b.aabb_min, sizeof(n.aabb_min));
    std::memcpy(n.aabb_max, b.aabb_max, sizeof(n.aabb_max));
    if (is_tri_leaf(b) || d == max_depth) {
        n.prim_count          = is_tri_leaf(b) ? NBVH_LEAF_FINAL : NBVH_LEAF_LOD(first_lod);
        n.left_or_prim_offset = NBVH_NO_NEURAL_CHILD;
        return;
    }
    const uint32_t left = (uint32_t)l->nodes.size();
    l->nodes.resize(left + 2);
    l->node_map.resize(left + 2);
    l->depth.resize(left + 2);
    l->nodes[ni].prim_count          = 0;
    l->nodes[ni].left_or_prim_offset = left;
    build_cut(l, b.left_or_prim_offset, left, d + 1, max_depth, first_lod);
    build_cut(l, b.left_or_prim_offset + 1, left + 1, d + 1, max_depth, first_lod);
}

void zero_words(void* p, uint64_t words, hipStream_t st)
{
    hip_check(train_zero(static_cast<float*>(p), words, st), "zero fill");
}

// Device node arrays sized for at least n nodes; the tree-cut data keeps its first `keep` entries.
void ensure_capacity(nbvh_learner* l, uint32_t n, uint32_t keep, hipStream_t st)
{
    if (n <= l->capacity) return;
    const uint32_t cap = std::max<uint32_t>({n, 2 * l->capacity, 256u});
    DevBuf tc;
    tc.alloc(l->ctx, (size_t)cap * sizeof(TreeCutNode));
    if (keep > 0)
        hip_check(hipMemcpyAsync(tc.ptr, l->d_tc.ptr, (size_t)keep * sizeof(TreeCutNode), hipMemcpyDeviceToDevice, st),
                  "tree-cut D2D");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::swap(l->d_tc.ctx, tc.ctx);
    std::swap(l->d_tc.ptr, tc.ptr);
    std::swap(l->d_tc.bytes, tc.bytes);
    tc.release();
    l->d_nodes.alloc(l->ctx, (size_t)cap * sizeof(nbvh_node));
    l->d_sums.alloc(l->ctx, (size_t)cap * sizeof(double));
    l->d_counts.alloc(l->ctx, (size_t)cap * sizeof(uint32_t));
    for (DevBuf* b : {&l->d_keys, &l->d_vals, &l->d_keys_sorted, &l->d_vals_sorted}) b->alloc(l->ctx, (size_t)cap * 4);
    l->sort_tmp_bytes = treecut_sort_scratch_bytes(cap);
    l->d_sort_tmp.alloc(l->ctx, std::max<size_t>(l->sort_tmp_bytes, 256));
    zero_words(l->d_sums.ptr, 2ull * cap, st);
    zero_words(l->d_counts.ptr, cap, st);
    l->capacity = cap;
}

void upload_nodes(nbvh_learner* l, hipStream_t st)
{
    hip_check(hipMemcpyAsync(l->d_nodes.ptr, l->nodes.data(), l->nodes.size() * sizeof(nbvh_node),
                             hipMemcpyHostToDevice, st),
              "nodes H2D");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    set_scene_nbvh(l->scene, l->nodes);
}

// split_neural_bvh_at (src/cuda/neural/neural_bvh_renderer_accel.cpp:256-392): the distinct
// candidates among the first min(#leaves, #candidates), in ascending node order; a candidate
// whose triangle-BVH node is a leaf, or that already has children, is skipped. The children
// (the triangle-BVH node's two children) are appended as leaves of the current LoD (LEAF_FINAL
// for triangle-BVH leaves) with fresh tree-cut data one level deeper.
void split(nbvh_learner* l, hipStream_t st)
{
    const size_t leaves = (l->nodes.size() + 1) / 2;
    const size_t m = std::min(leaves, l->split_idx.size());
    const std::set<uint32_t> unique(l->split_idx.begin(), l->split_idx.begin() + (long)m);
    const std::vector<nbvh_node>& bvh = l->scene->bvh;
    const uint32_t n0 = (uint32_t)l->nodes.size();
    std::vector<TreeCutNode> fresh;
    for (uint32_t s : unique) {
        const nbvh_node& parent = bvh[l->node_map[s]];
        if (is_tri_leaf(parent) || l->nodes[s].left_or_prim_offset != NBVH_NO_NEURAL_CHILD) continue;
        const uint32_t left = (uint32_t)l->nodes.size();
        l->nodes.resize(left + 2);
        l->node_map.push_back(parent.left_or_prim_offset);
        l->node_map.push_back(parent.left_or_prim_offset + 1);
        l->depth.push_back(l->depth[s] + 1);
        l->depth.push_back(l->depth[s] + 1);
        nbvh_node& n = l->nodes[s];
        if (n.prim_count == NBVH_LEAF_LOD(l->current_lod)) n.prim_count = 0;
        n.left_or_prim_offset = left;
        for (uint32_t c = 0; c < 2; c++) {
            const nbvh_node& tb = bvh[parent.left_or_prim_offset + c];
            nbvh_node& ch = l->nodes[left + c];
            std::memcpy(ch.aabb_min, tb.aabb_min, sizeof(ch.aabb_min));
            std::memcpy(ch.aabb_max, tb.aabb_max, sizeof(ch.aabb_max));
            ch.prim_count          = is_tri_leaf(tb) ? NBVH_LEAF_FINAL : NBVH_LEAF_LOD(l->current_lod);
            ch.left_or_prim_offset = NBVH_NO_NEURAL_CHILD;
            fresh.push_back(TreeCutNode{0.f, 1u, l->depth[s] + 1, 0u});
        }
        l->total_splits++;
    }
    if (fresh.empty()) return;
    ensure_capacity(l, (uint32_t)l->nodes.size(), n0, st);
    hip_check(hipMemcpyAsync(static_cast<TreeCutNode*>(l->d_tc.ptr) + n0, fresh.data(),
                             fresh.size() * sizeof(TreeCutNode), hipMemcpyHostToDevice, st),
              "tree-cut H2D");
    upload_nodes(l, st);
}

// assign_lod_to_current_neural_leaf_nodes (src/cuda/neural/neural_bvh_renderer_accel.cpp:395-440).
void assign_lod(nbvh_learner* l, hipStream_t st)
{
    const uint32_t lod = l->current_lod;
    if (lod != NBVH_NUM_LODS - 1) {
        for (nbvh_node& n : l->nodes)
            if (n.left_or_prim_offset == NBVH_NO_NEURAL_CHILD)
                n.prim_count = n.prim_count == NBVH_LEAF_FINAL ? NBVH_LEAF_FINAL : NBVH_LEAF_LOD(lod);
        upload_nodes(l, st);
    }
    l->all_assigned = lod == 0;
    if (!l->all_assigned)
        l->current_lod--;
    else
        l->tree_cut = false;
}

void sample(nbvh_learner* l, uint32_t step, hipStream_t st)
{
    const nbvh_scene* s = l->scene;
    SampleArgs a{};
    a.nodes   = l->d_nodes.as<nbvh_node>();
    a.n_nodes = (uint32_t)l->nodes.size();
    a.tri = TriArgs{s->d_verts.as<float>(), s->d_tris.as<uint32_t>(), s->d_bvh.as<nbvh_node>(), s->d_prims.as<uint32_t>(),
                    0u};
    for (int c = 0; c < 3; c++) {
        a.box_min[c]    = s->bbox_min[c];
        a.box_ext[c]    = s->bbox_max[c] - s->bbox_min[c];
        a.enc_center[c] = s->args.enc_center[c];
        a.enc_extent[c] = s->args.enc_extent[c];
    }
    for (int k = 0; k < 4; k++) a.seg_t[k] = s->args.seg_t[k];
    a.n_points         = s->args.n_points;
    a.scaling          = l->cfg.sampling_bbox_scaling;
    a.fixed_dilation   = l->cfg.fixed_node_dilation;
    a.extent_dilation  = l->cfg.node_extent_dilation;
    a.mint_eps         = l->cfg.learning_mint_epsilon;
    a.maxt_eps         = l->cfg.learning_maxt_epsilon;
    a.current_lod      = l->current_lod;
    a.survival         = l->cfg.training_node_survival ? 1u : 0u;
    a.survival_scale   = l->cfg.survival_probability_scale;
    a.min_survival     = l->cfg.min_rr_survival_prob;
    a.current_max_loss = step == 0 ? 0.f : l->max_node_loss;
    a.tree_cut         = l->d_tc.as<TreeCutNode>();
    a.n_sub_batch      = 1u << l->cfg.log2_bvh_max_traversals;
    a.sub_batch_size   = l->cfg.batch_size / a.n_sub_batch;
    a.batch_size       = l->cfg.batch_size;
    a.step             = step;
    a.inputs           = l->d_inputs.as<float>();
    a.targets          = l->d_targets.as<float>();
    a.node_idx         = l->d_node_idx.as<uint32_t>();
    a.failures         = l->d_failures.as<uint32_t>();
    hip_check(learn_sample(a, st), "sampler launch");
}

// The tree-cut part of a step (the reference's loop after training_step).
void update(nbvh_learner* l, const float* d_sample_loss, hipStream_t st)
{
    bool should_split = false;
    if (l->tree_cut) {
        const int n = l->sched.splits((int)l->step);
        if ((uint32_t)n != l->current_num_splits) l->current_num_splits = (uint32_t)std::max(1, n);
        should_split = n > 0;
    }
    const uint32_t n_nodes = (uint32_t)l->nodes.size();
    hip_check(treecut_update(l->d_node_idx.as<uint32_t>(), d_sample_loss, l->cfg.batch_size, l->d_nodes.as<nbvh_node>(),
                             n_nodes, l->d_tc.as<TreeCutNode>(), static_cast<double*>(l->d_sums.ptr),
                             l->d_counts.as<uint32_t>(), l->d_keys.as<float>(), l->d_vals.as<uint32_t>(),
                             l->d_keys_sorted.as<float>(), l->d_vals_sorted.as<uint32_t>(), l->d_sort_tmp.ptr,
                             l->sort_tmp_bytes, st),
              "tree-cut update");
    // Candidates: the leaves of the current_num_splits largest distinct positive heuristics.
    const uint32_t k = l->current_num_splits;
    l->split_idx.assign(k, 0u);
    uint32_t m = std::min<uint32_t>(n_nodes, std::max<uint32_t>(256u, 4u * k));
    std::vector<float> keys;
    std::vector<uint32_t> vals;
    uint32_t found = 0;
    float first = 0.f;
    while (true) {
        keys.resize(m);
        vals.resize(m);
        hip_check(hipMemcpyAsync(keys.data(), l->d_keys_sorted.ptr, m * 4, hipMemcpyDeviceToHost, st), "keys D2H");
        hip_check(hipMemcpyAsync(vals.data(), l->d_vals_sorted.ptr, m * 4, hipMemcpyDeviceToHost, st), "vals D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        found = 0;
        float prev = 0.f;
        uint32_t j = 0;
        for (; j < m && found < k && keys[j] > 0.f; j++) {
            if (found > 0 && keys[j] == prev) continue;
            l->split_idx[found++] = vals[j];
            prev = keys[j];
        }
        first = m > 0 && keys[0] > 0.f ? keys[0] : 0.f;
        if (found == k || j < m || m == n_nodes) break;
        m = n_nodes;  // ties filled the window: read the whole order
    }
    l->max_node_loss = first;
    if (l->tree_cut) {
        if (should_split) split(l, st);
        if (l->cfg.num_lods_learned > 1 && !l->all_assigned) {
            if (l->sched.should_assign_new_lod((int)l->step, (int)l->current_lod, (int)l->cfg.num_lods_learned,
                                               (int)NBVH_NUM_LODS))
                assign_lod(l, st);
            if (l->all_assigned) l->current_num_splits = 1;
        }
    }
    l->step++;
}

void check_config(const nbvh_learner_config& c)
{
    require(c.batch_size >= 1 && c.batch_size <= (1u << 24), "batch_size must be 1..2^24");
    require(c.log2_bvh_max_traversals <= 16, "log2_bvh_max_traversals must be 0..16");
    require(c.batch_size % (1u << c.log2_bvh_max_traversals) == 0,
            "batch_size must be a multiple of 2^log2_bvh_max_traversals");
    require(c.sampling_bbox_scaling > 0.f && c.fixed_node_dilation >= 0.f && c.node_extent_dilation >= 0.f &&
                c.learning_mint_epsilon >= 0.f && c.learning_maxt_epsilon >= 0.f,
            "sampler parameters out of range");
    require(c.num_lods_learned >= 1 && c.num_lods_learned <= NBVH_NUM_LODS, "num_lods_learned must be 1..7");
    require(c.lod_decision <= 1, "lod_decision must be 0 or 1");
    // An interval scaling below 1 can truncate the schedule's interval to 0 (a modulo by zero in
    // the reference's scheduler).
    require(c.splitting_interval > 0 && c.split_scaling > 0.f && c.split_interval_scaling >= 1.f &&
                c.start_split_num >= 0 && c.last_split_iteration >= 0,
            "split schedule parameters out of range");
    if (c.max_training_bounces > 1)
        fail(NBVH_ERR_UNSUPPORTED, "training bounces (max_training_bounces > 1) are not supported");
}

}  // namespace

NBVH_API int32_t nbvh_learner_config_defaults(nbvh_learner_config* c)
{
    return guarded(nullptr, [&] {
        require(c != nullptr, "NULL argument");
        std::memset(c, 0, sizeof(*c));
        c->batch_size                 = 1u << 18;
        c->log2_bvh_max_traversals    = 0;
        c->sampling_bbox_scaling      = 1.5f;
        c->fixed_node_dilation        = 1e-5f;
        c->node_extent_dilation       = 1e-4f;
        c->learning_mint_epsilon      = 1e-3f;
        c->learning_maxt_epsilon      = 1e-3f;
        c->max_training_bounces       = 1;
        c->training_node_survival     = 1;
        c->survival_probability_scale = 1.f;
        c->min_rr_survival_prob       = 0.005f;
        c->tree_cut_learning          = 1;
        c->neural_bvh_min_depth       = 0;
        c->num_lods_learned           = 1;
        c->max_training_steps         = 10000;
        c->lod_decision               = 0;
        c->splitting_interval         = 50;
        c->last_split_iteration       = 2000;
        c->start_split_num            = 1;
        c->split_scaling              = 2.f;
        c->split_interval_scaling     = 2.f;
    });
}

NBVH_API int32_t nbvh_learner_create(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_learner_config* cfg,
                                     nbvh_learner** out)
{
    return guarded(ctx, [&] {
        require(ctx && scene && cfg && out, "NULL argument");
        require(scene->ctx == ctx, "scene belongs to another context");
        *out = nullptr;
        check_config(*cfg);
        if (scene->bvh.empty()) fail(NBVH_ERR_STATE, "tree-cut learning needs the scene's triangles");
        if (!scene->has_net) fail(NBVH_ERR_STATE, "tree-cut learning needs the scene's network (its encoding)");
        if (scene->cfg.segment_distribution == NBVH_SEGMENT_UNIFORM)
            fail(NBVH_ERR_UNSUPPORTED, "tree-cut learning with the uniform segment distribution");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto l = std::make_unique<nbvh_learner>();
        l->ctx   = ctx;
        l->scene = scene;
        l->cfg   = *cfg;
        l->sched.interval         = cfg->splitting_interval;
        l->sched.last_iteration   = cfg->last_split_iteration;
        l->sched.start_num        = cfg->start_split_num;
        l->sched.scaling          = cfg->split_scaling;
        l->sched.interval_scaling = cfg->split_interval_scaling;
        l->sched.lod_decision     = cfg->lod_decision;
        // build_and_upload_neural_bvh + reinitialise_neural_bvh_learning_states: the cut at
        // min depth with every cut leaf at the coarsest LoD, which is then the LoD being learned.
        l->nodes.resize(1);
        l->node_map.resize(1);
        l->depth.resize(1);
        build_cut(l.get(), 0, 0, 0, cfg->neural_bvh_min_depth, NBVH_NUM_LODS - 1);
        l->current_lod  = NBVH_NUM_LODS - 1;
        l->all_assigned = false;
        l->tree_cut     = cfg->tree_cut_learning != 0;
        l->current_num_splits = (uint32_t)std::max(1, cfg->start_split_num);
        l->sched.update_stats_and_reset((int)l->nodes.size(), (int)cfg->max_training_steps);
        const hipStream_t st = nullptr;
        ensure_capacity(l.get(), (uint32_t)l->nodes.size(), 0, st);
        std::vector<TreeCutNode> tc(l->nodes.size());
        for (size_t i = 0; i < tc.size(); i++) tc[i] = TreeCutNode{0.f, 1u, l->depth[i], 0u};
        hip_check(hipMemcpy(l->d_tc.ptr, tc.data(), tc.size() * sizeof(TreeCutNode), hipMemcpyHostToDevice),
                  "tree-cut H2D");
        const size_t b = cfg->batch_size;
        l->d_inputs.alloc(ctx, b * n_inputs(l.get()) * sizeof(float));
        l->d_targets.alloc(ctx, b * 8 * sizeof(float));
        l->d_node_idx.alloc(ctx, b * sizeof(uint32_t));
        l->d_sample_loss.alloc(ctx, b * sizeof(float));
        l->d_failures.alloc(ctx, 256);
        // Slots that no sample ever reached hold zeros (the reference's are uninitialised).
        zero_words(l->d_inputs.ptr, b * n_inputs(l.get()), st);
        zero_words(l->d_targets.ptr, b * 8, st);
        zero_words(l->d_node_idx.ptr, b, st);
        zero_words(l->d_sample_loss.ptr, b, st);
        zero_words(l->d_failures.ptr, 1, st);
        upload_nodes(l.get(), st);
        *out = l.release();
    });
}

NBVH_API int32_t nbvh_learner_destroy(nbvh_learner* l)
{
    if (!l) return NBVH_OK;
    (void)hipSetDevice(l->ctx->device);
    (void)hipDeviceSynchronize();
    delete l;
    return NBVH_OK;
}

NBVH_API int32_t nbvh_learner_sample(nbvh_learner* l, uint32_t step, void* stream)
{
    return guarded(l ? l->ctx : nullptr, [&] {
        require(l != nullptr, "NULL learner");
        hip_check(hipSetDevice(l->ctx->device), "hipSetDevice");
        sample(l, step, (hipStream_t)stream);
    });
}

NBVH_API int32_t nbvh_learner_update(nbvh_learner* l, const float* d_sample_loss, void* stream)
{
    return guarded(l ? l->ctx : nullptr, [&] {
        require(l && d_sample_loss, "NULL argument");
        hip_check(hipSetDevice(l->ctx->device), "hipSetDevice");
        update(l, d_sample_loss, (hipStream_t)stream);
    });
}

NBVH_API int32_t nbvh_learner_step(nbvh_learner* l, nbvh_trainer* t, float* loss, void* stream)
{
    return guarded(l ? l->ctx : nullptr, [&] {
        require(l && t, "NULL argument");
        if (trainer_scene(t) != l->scene) fail(NBVH_ERR_INVALID_ARGUMENT, "the trainer trains another scene");
        const hipStream_t st = (hipStream_t)stream;
        hip_check(hipSetDevice(l->ctx->device), "hipSetDevice");
        sample(l, l->step, st);
        float value = 0.f;
        trainer_step(t, l->d_inputs.as<float>(), l->d_targets.as<float>(), l->cfg.batch_size, 0u,
                     l->d_sample_loss.as<float>(), &value, st);
        l->last_loss = value;
        update(l, l->d_sample_loss.as<float>(), st);
        if (loss) *loss = value;
    });
}

NBVH_API int32_t nbvh_learner_get_state(const nbvh_learner* l, nbvh_learner_state* s)
{
    return guarded(l ? l->ctx : nullptr, [&] {
        require(l && s, "NULL argument");
        std::memset(s, 0, sizeof(*s));
        uint32_t failures = 0;
        hip_check(hipSetDevice(l->ctx->device), "hipSetDevice");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpy(&failures, l->d_failures.ptr, 4, hipMemcpyDeviceToHost), "failures D2H");
        s->step                  = l->step;
        s->n_nodes               = (uint32_t)l->nodes.size();
        s->current_learned_lod   = l->current_lod;
        s->all_lods_assigned     = l->all_assigned ? 1u : 0u;
        s->tree_cut_learning     = l->tree_cut ? 1u : 0u;
        s->current_num_splits    = l->current_num_splits;
        s->total_splits          = l->total_splits;
        s->sample_failures       = failures;
        s->max_node_loss         = l->max_node_loss;
        s->last_loss             = l->last_loss;
        s->expected_total_splits = (uint32_t)l->sched.stats_total_splits;
    });
}

NBVH_API int32_t nbvh_learner_get(nbvh_learner* l, uint32_t what, void* out, uint64_t capacity, uint64_t* count)
{
    return guarded(l ? l->ctx : nullptr, [&] {
        require(l != nullptr, "NULL learner");
        const uint64_t b = l->cfg.batch_size, nn = l->nodes.size();
        const void* src = nullptr;
        const void* host = nullptr;
        uint64_t n = 0, elem = 4;
        switch (what) {
        case NBVH_LEARN_INPUTS: src = l->d_inputs.ptr; n = b * n_inputs(l); break;
        case NBVH_LEARN_TARGETS: src = l->d_targets.ptr; n = b * 8; break;
        case NBVH_LEARN_NODES: src = l->d_node_idx.ptr; n = b; break;
        case NBVH_LEARN_TREE_CUT: src = l->d_tc.ptr; n = nn; elem = sizeof(TreeCutNode); break;
        case NBVH_LEARN_SPLITS: host = l->split_idx.data(); n = l->split_idx.size(); break;
        case NBVH_LEARN_NBVH: host = l->nodes.data(); n = nn; elem = sizeof(nbvh_node); break;
        default: fail(NBVH_ERR_INVALID_ARGUMENT, "unknown learner array");
        }
        if (count) *count = n;
        if (!out) return;
        if (capacity < n) fail(NBVH_ERR_CAPACITY, "output buffer too small");
        if (host) {
            if (n) std::memcpy(out, host, n * elem);
            return;
        }
        hip_check(hipSetDevice(l->ctx->device), "hipSetDevice");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (n) hip_check(hipMemcpy(out, src, n * elem, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
}
