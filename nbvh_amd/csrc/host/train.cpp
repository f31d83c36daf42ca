// train.cpp -- the trainer object behind nbvh_trainer_*: a device copy of a scene's network in
// FP32, its gradients and Adam moments, and the step loop (kernels/train.hip).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>

#include "../kernels/launch.h"
#include "context.h"

using namespace nbvh;

struct nbvh_trainer {
    nbvh_ctx* ctx     = nullptr;
    nbvh_scene* scene = nullptr;
    nbvh_train_config cfg{};
    uint64_t n_table = 0, n_mlp = 0;
    uint32_t step    = 0;
    uint64_t net_generation = 0;  // the scene's network generation the buffers were sized for
    NetArgs net{};                // the scene's network arguments at creation (shape and levels)
    DevBuf tables, mlp, g_tables, g_mlp, m_tables, v_tables, m_mlp, v_mlp, n_tables, n_mlp_steps, act, bins, loss;
};

namespace {

// The training kernels implement the shapes of the MFMA query: 96 inputs (P*L*F), 4 hidden
// layers of 64, 8 outputs.
bool trainable(const nbvh_network_config& c)
{
    const bool shape = (c.n_points == 3 && c.grid.n_levels == 16 && c.grid.n_features_per_level == 2) ||
                       (c.n_points == 3 && c.grid.n_levels == 8 && c.grid.n_features_per_level == 4);
    // The loss trains the LocalT depth through the sigmoid output (the reference's defaults).
    return shape && c.n_neurons == 64 && c.n_hidden_layers == 4 && c.n_outputs == 8 &&
           c.depth_mode == NBVH_DEPTH_LOCAL_T && c.depth_linear_activation == 0;
}

// Learning-rate factor of update number `step` (from 1) under the exponential decay: decay_base to
// the number of earlier updates s >= decay_start with (s - decay_start) % decay_interval == 0.
float decay_factor(const nbvh_train_config& c, uint64_t step)
{
    if (c.decay_interval == 0 || step <= 1 || step - 1 < c.decay_start) return 1.f;
    // s = decay_start, decay_start + interval, ... <= step - 1; s = 0 is not an update.
    const uint64_t k = (step - 1 - c.decay_start) / c.decay_interval + (c.decay_start > 0 ? 1 : 0);
    return (float)std::pow((double)c.decay_base, (double)k);
}

// The trainer's buffers are sized for the network the scene held at nbvh_trainer_create; a later
// nbvh_scene_load_network / nbvh_scene_init_network replaces it (possibly with other table sizes),
// after which the trainer refuses to run or commit.
void check_generation(const nbvh_trainer* t)
{
    if (t->scene->net_generation != t->net_generation)
        fail(NBVH_ERR_STATE, "the scene's network was replaced after the trainer was created");
}

TrainArgs train_args(nbvh_trainer* t)
{
    TrainArgs a{};
    a.net         = t->net;
    a.tables      = t->tables.as<float>();
    a.mlp         = t->mlp.as<float>();
    a.act         = t->act.as<float>();
    a.act_stride  = train_act_stride(t->cfg.max_batch);
    a.grad_tables = t->g_tables.as<float>();
    a.grad_mlp    = t->g_mlp.as<float>();
    a.loss        = t->loss.as<float>();
    a.bins        = t->bins.ptr;
    return a;
}

}  // namespace

NBVH_API int32_t nbvh_trainer_create(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_train_config* cfg,
                                     nbvh_trainer** out)
{
    return guarded(ctx, [&] {
        require(ctx && scene && cfg && out, "NULL argument");
        require(scene->ctx == ctx, "scene belongs to another context");
        *out = nullptr;
        if (!scene->has_net) fail(NBVH_ERR_STATE, "training a scene without network weights");
        if (!trainable(scene->cfg)) fail(NBVH_ERR_UNSUPPORTED, "no training kernels for this network shape");
        require(cfg->max_batch >= 1 && cfg->max_batch <= (1u << 24), "max_batch must be 1..2^24");
        // the table gradient's corner terms of a step are indexed with 32 bits
        require((uint64_t)train_act_stride(cfg->max_batch) * scene->cfg.n_points * scene->cfg.grid.n_levels * 8u <
                    (1ull << 32),
                "max_batch too large for this network: max_batch * n_points * n_levels * 8 must be below 2^32");
        require(cfg->decay_interval == 0 || (std::isfinite(cfg->decay_base) && cfg->decay_base > 0.f),
                "decay_base must be finite and > 0");
        require(cfg->learning_rate > 0.f && cfg->beta1 >= 0.f && cfg->beta1 < 1.f && cfg->beta2 >= 0.f &&
                    cfg->beta2 < 1.f && cfg->epsilon >= 0.f && cfg->l2_reg >= 0.f,
                "optimiser parameters out of range");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto t = std::make_unique<nbvh_trainer>();
        t->ctx = ctx;
        t->scene = scene;
        t->cfg = *cfg;
        t->n_table = scene->tables.size();
        t->n_mlp = scene->mlp.size();
        t->net_generation = scene->net_generation;
        t->net = scene->args;
        t->tables.upload(ctx, scene->tables.data(), t->n_table);
        t->mlp.upload(ctx, scene->mlp.data(), t->n_mlp);
        for (DevBuf* b : {&t->g_tables, &t->m_tables, &t->v_tables}) b->alloc(ctx, t->n_table * sizeof(float));
        for (DevBuf* b : {&t->g_mlp, &t->m_mlp, &t->v_mlp}) b->alloc(ctx, t->n_mlp * sizeof(float));
        t->n_tables.alloc(ctx, t->n_table * sizeof(uint32_t));
        t->n_mlp_steps.alloc(ctx, t->n_mlp * sizeof(uint32_t));
        t->act.alloc(ctx, train_act_bytes(cfg->max_batch));
        t->bins.alloc(ctx, train_bin_bytes(cfg->max_batch, scene->cfg.n_points, scene->cfg.grid.n_levels,
                                           scene->cfg.grid.n_features_per_level));
        t->loss.alloc(ctx, 256);
        hip_check(train_zero(t->m_tables.as<float>(), t->n_table, nullptr), "zero");
        hip_check(train_zero(t->v_tables.as<float>(), t->n_table, nullptr), "zero");
        hip_check(train_zero(t->m_mlp.as<float>(), t->n_mlp, nullptr), "zero");
        hip_check(train_zero(t->v_mlp.as<float>(), t->n_mlp, nullptr), "zero");
        hip_check(train_zero(t->n_tables.as<float>(), t->n_table, nullptr), "zero");  // update counts
        hip_check(train_zero(t->n_mlp_steps.as<float>(), t->n_mlp, nullptr), "zero");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        *out = t.release();
    });
}

NBVH_API int32_t nbvh_trainer_destroy(nbvh_trainer* t)
{
    if (!t) return NBVH_OK;
    (void)hipSetDevice(t->ctx->device);
    (void)hipDeviceSynchronize();
    delete t;
    return NBVH_OK;
}

namespace nbvh {

nbvh_scene* trainer_scene(const nbvh_trainer* t) { return t->scene; }

void trainer_step(nbvh_trainer* t, const float* d_inputs, const float* d_targets, uint32_t n, uint32_t flags,
                  float* d_sample_loss, float* loss, hipStream_t st)
{
    require(t != nullptr, "NULL trainer");
    require((flags & ~(uint32_t)NBVH_TRAIN_GRADIENTS_ONLY) == 0u, "unknown training flags");
    require(n <= t->cfg.max_batch, "batch larger than the trainer's max_batch");
    if (n > 0) require(d_inputs && d_targets, "NULL buffer");
    check_generation(t);
    hip_check(hipSetDevice(t->ctx->device), "hipSetDevice");
    hip_check(train_zero(t->g_mlp.as<float>(), t->n_mlp, st), "zero");
    hip_check(train_zero(t->loss.as<float>(), 1, st), "zero");
    bool supported = false;
    TrainArgs args = train_args(t);
    args.sample_loss = d_sample_loss;
    hip_check(train_gradients(args, d_inputs, d_targets, n, st, &supported), "training kernels");
    if (!supported) fail(NBVH_ERR_UNSUPPORTED, "no training kernels for this network shape");
    if ((flags & NBVH_TRAIN_GRADIENTS_ONLY) == 0u && n > 0) {
        t->step++;
        AdamArgs h{};
        h.lr        = t->cfg.learning_rate * decay_factor(t->cfg, t->step);
        h.beta1     = t->cfg.beta1;
        h.beta2     = t->cfg.beta2;
        h.epsilon   = t->cfg.epsilon;
        h.l2_reg    = t->cfg.l2_reg;
        hip_check(train_adam(t->tables.as<float>(), t->g_tables.as<float>(), t->m_tables.as<float>(),
                             t->v_tables.as<float>(), t->n_tables.as<uint32_t>(), t->n_table, false, h, st),
                  "adam tables");
        hip_check(train_adam(t->mlp.as<float>(), t->g_mlp.as<float>(), t->m_mlp.as<float>(), t->v_mlp.as<float>(),
                             t->n_mlp_steps.as<uint32_t>(), t->n_mlp, true, h, st),
                  "adam mlp");
    }
    if (loss) {
        hip_check(hipMemcpyAsync(loss, t->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, st), "loss D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
}

}  // namespace nbvh

NBVH_API int32_t nbvh_trainer_step(nbvh_trainer* t, const float* d_inputs, const float* d_targets, uint32_t n,
                                   uint32_t flags, float* loss, void* stream)
{
    return guarded(t ? t->ctx : nullptr,
                   [&] { trainer_step(t, d_inputs, d_targets, n, flags, nullptr, loss, (hipStream_t)stream); });
}

NBVH_API int32_t nbvh_trainer_get(nbvh_trainer* t, uint32_t what, float* out, uint64_t capacity, uint64_t* count)
{
    return guarded(t ? t->ctx : nullptr, [&] {
        require(t != nullptr, "NULL trainer");
        const DevBuf* src = nullptr;
        uint64_t n = 0;
        switch (what) {
        case NBVH_TRAIN_TABLES: src = &t->tables; n = t->n_table; break;
        case NBVH_TRAIN_MLP: src = &t->mlp; n = t->n_mlp; break;
        case NBVH_TRAIN_GRAD_TABLES: src = &t->g_tables; n = t->n_table; break;
        case NBVH_TRAIN_GRAD_MLP: src = &t->g_mlp; n = t->n_mlp; break;
        default: fail(NBVH_ERR_INVALID_ARGUMENT, "unknown trainer array");
        }
        if (count) *count = n;
        if (!out) return;
        if (capacity < n) fail(NBVH_ERR_CAPACITY, "output buffer too small");
        hip_check(hipSetDevice(t->ctx->device), "hipSetDevice");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpy(out, src->ptr, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
}

NBVH_API int32_t nbvh_trainer_commit(nbvh_trainer* t)
{
    return guarded(t ? t->ctx : nullptr, [&] {
        require(t != nullptr, "NULL trainer");
        check_generation(t);
        nbvh_scene* s = t->scene;
        hip_check(hipSetDevice(t->ctx->device), "hipSetDevice");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpy(s->tables.data(), t->tables.ptr, t->n_table * sizeof(float), hipMemcpyDeviceToHost),
                  "tables D2H");
        hip_check(hipMemcpy(s->mlp.data(), t->mlp.ptr, t->n_mlp * sizeof(float), hipMemcpyDeviceToHost), "mlp D2H");
        upload_network(s);
        refresh_args(s);
    });
}
