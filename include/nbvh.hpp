/*
 * nbvh.hpp -- C++ host interface over the C ABI of libnbvh_hip.so (nbvh.h), header-only.
 *
 * Mirrors the reference's host-side classes for the neural ray query so that C++ callers keep
 * their shape:
 *   ntwr::neural::NeuralBVHModule / NeuralPathModule   include/cuda/neural/neural_bvh_module.h:144-241,
 *                                                      include/cuda/neural/path/path_module.h:256-350
 *     neural_bvh_traversal(...)                        src/cuda/neural/path/path_module.cu:54-137
 *   NeuralBVHRenderer::load_neural_bvh                 src/cuda/neural/neural_bvh_renderer_accel.cpp:208-254
 *   NeuralBVHRenderer::load_network_model              src/cuda/neural/neural_bvh_renderer_init.cu:394-404
 * Failures throw nbvh::Error (a std::runtime_error carrying the status), as the reference's
 * CUDA_CHECK throws std::runtime_error (include/ntwr.h:30-39). Nothing here touches the GPU
 * directly: every call goes through the C ABI.
 */
#ifndef NBVH_HPP_
#define NBVH_HPP_

#include <cstdint>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nbvh.h"

namespace nbvh {

class Error : public std::runtime_error {
public:
    Error(int32_t status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
    int32_t status() const { return status_; }

private:
    int32_t status_;
};

inline void check(int32_t status, const nbvh_ctx* ctx = nullptr)
{
    if (status == NBVH_OK) return;
    const char* m = nbvh_last_error(ctx);
    throw Error(status, std::string("nbvh: ") + nbvh_status_string(status) + (m && *m ? std::string(": ") + m : ""));
}

inline nbvh_network_config network_config(uint32_t n_levels = 16, uint32_t n_features_per_level = 2,
                                          uint32_t log2_hashmap_size = 19, uint32_t base_resolution = 16,
                                          float per_level_scale = 1.5f)
{
    nbvh_network_config c{};
    c.grid                 = {n_levels, n_features_per_level, log2_hashmap_size, base_resolution, per_level_scale};
    c.n_points             = 3;
    c.segment_distribution = NBVH_SEGMENT_EQUIDISTANT_EXTERIOR;
    c.n_neurons            = 64;
    c.n_hidden_layers      = 4;
    c.n_outputs            = 8;
    return c;
}

struct Mesh {
    std::vector<float> verts;     // xyz
    std::vector<uint32_t> tris;   // 3 indices per triangle
    uint32_t n_verts() const { return (uint32_t)(verts.size() / 3); }
    uint32_t n_tris() const { return (uint32_t)(tris.size() / 3); }
};

inline Mesh icosphere(uint32_t subdivisions, uint64_t seed = 0, float amplitude = 0.1f, const float center[3] = nullptr,
                      float radius = 1.f)
{
    uint32_t nv = 0, nt = 0;
    check(nbvh_mesh_icosphere(subdivisions, seed, amplitude, center, radius, nullptr, &nv, nullptr, &nt));
    Mesh m;
    m.verts.resize((size_t)nv * 3);
    m.tris.resize((size_t)nt * 3);
    check(nbvh_mesh_icosphere(subdivisions, seed, amplitude, center, radius, m.verts.data(), &nv, m.tris.data(), &nt));
    return m;
}

inline Mesh geosphere(uint32_t frequency, uint64_t seed = 0, float amplitude = 0.1f, const float center[3] = nullptr,
                      float radius = 1.f)
{
    uint32_t nv = 0, nt = 0;
    check(nbvh_mesh_geosphere(frequency, seed, amplitude, center, radius, nullptr, &nv, nullptr, &nt));
    Mesh m;
    m.verts.resize((size_t)nv * 3);
    m.tris.resize((size_t)nt * 3);
    check(nbvh_mesh_geosphere(frequency, seed, amplitude, center, radius, m.verts.data(), &nv, m.tris.data(), &nt));
    return m;
}

inline nbvh_camera camera_look_at(const float eye[3], const float target[3], const float up[3], float fovy_degrees,
                                  uint32_t width, uint32_t height)
{
    nbvh_camera c{};
    check(nbvh_camera_look_at(eye, target, up, fovy_degrees, width, height, &c));
    return c;
}

// One GPU with an HBM budget (nbvh_create / nbvh_destroy).
class Context {
public:
    explicit Context(int32_t device = 0, uint64_t hbm_budget_bytes = 0)
    {
        nbvh_create_info info{device, 0, hbm_budget_bytes};
        check(nbvh_create(&info, &ctx_));
    }
    ~Context() { nbvh_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    nbvh_ctx* get() const { return ctx_; }

    // Primary rays into device memory (d_rays: width*height records), on `stream` (hipStream_t).
    void raygen(const nbvh_camera& cam, uint32_t width, uint32_t height, nbvh_ray* d_rays, uint32_t sample = 0,
                uint32_t frame = 0, void* stream = nullptr)
    {
        check(nbvh_raygen(ctx_, &cam, width, height, sample, frame, d_rays, stream), ctx_);
    }
    std::pair<uint64_t, uint64_t> memory_usage() const
    {
        uint64_t used = 0, budget = 0;
        check(nbvh_memory_usage(ctx_, &used, &budget), ctx_);
        return {used, budget};
    }

private:
    nbvh_ctx* ctx_ = nullptr;
};

// Device-resident scene: triangle BVH (fallback), N-BVH, encoding box, network weights.
class Scene {
public:
    explicit Scene(Context& ctx) : ctx_(ctx.get()) { check(nbvh_scene_create_empty(ctx_, &s_), ctx_); }
    Scene(Context& ctx, const Mesh& m) : ctx_(ctx.get())
    {
        check(nbvh_scene_create(ctx_, m.verts.data(), m.n_verts(), m.tris.data(), m.n_tris(), &s_), ctx_);
    }
    ~Scene() { nbvh_scene_destroy(s_); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    nbvh_scene* get() const { return s_; }
    nbvh_ctx* ctx() const { return ctx_; }

    void build_neural_bvh(uint32_t n_leaves) { check(nbvh_scene_build_nbvh(s_, n_leaves), ctx_); }
    // NeuralBVHRenderer::load_neural_bvh: a .nbvh file written by the reference.
    void load_neural_bvh(const std::string& path)
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw Error(NBVH_ERR_INVALID_ARGUMENT, "cannot open " + path);
        std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        check(nbvh_scene_load_nbvh(s_, bytes.data(), bytes.size()), ctx_);
    }
    std::vector<nbvh_node> neural_bvh() const
    {
        uint32_t n = 0;
        check(nbvh_scene_get_nbvh(s_, nullptr, 0, &n), ctx_);
        std::vector<nbvh_node> out(n);
        check(nbvh_scene_get_nbvh(s_, out.data(), n, &n), ctx_);
        return out;
    }
    void set_bbox(const float mn[3], const float mx[3]) { check(nbvh_scene_set_bbox(s_, mn, mx), ctx_); }
    // NeuralBVHRenderer::load_network_model: tables in tcnn parameter order, MLP row-major per layer.
    void load_network_model(const nbvh_network_config& cfg, const std::vector<float>& tables,
                            const std::vector<float>& mlp)
    {
        check(nbvh_scene_load_network(s_, &cfg, tables.data(), tables.size(), mlp.data(), mlp.size()), ctx_);
    }
    void init_network(const nbvh_network_config& cfg, uint64_t seed, float table_range = 1e-4f)
    {
        check(nbvh_scene_init_network(s_, &cfg, seed, table_range), ctx_);
    }

private:
    nbvh_ctx* ctx_ = nullptr;
    nbvh_scene* s_ = nullptr;
};

struct QueryOptions {
    uint32_t kind                = NBVH_QUERY_NEURAL_CLOSEST;
    uint32_t precision           = NBVH_PREC_FP16;
    uint32_t lod                 = 0;
    uint32_t max_traversal_steps = 0;     // 0: 1000 (m_max_traversal_steps)
    float visibility_threshold   = 0.f;   // 0: 0.5
    uint32_t image_width         = 0;     // scheduling hint for image-ordered rays
    uint32_t flags               = 0;     // nbvh_query_flags (NBVH_QUERY_SORT_RAYS)
    uint32_t segment_seed        = 0;     // rays' LCG streams under NBVH_SEGMENT_UNIFORM
    nbvh_query_params params() const
    {
        nbvh_query_params p{};
        p.flags                = flags;
        p.segment_seed         = segment_seed;
        p.kind                 = kind;
        p.precision            = precision;
        p.lod                  = lod;
        p.max_traversal_steps  = max_traversal_steps;
        p.visibility_threshold = visibility_threshold;
        p.image_width          = image_width;
        return p;
    }
};

// Inference side of NeuralBVHModule / NeuralPathModule: a ray batch in, one hit record per ray out.
class NeuralBVHModule {
public:
    NeuralBVHModule(Context& ctx, Scene& scene) : ctx_(ctx.get()), scene_(scene.get()) {}

    // Device buffers, asynchronous on `stream` (the whole leaf loop of neural_bvh_traversal).
    void neural_bvh_traversal(const nbvh_ray* d_rays, uint32_t n_rays, nbvh_hit* d_hits, const QueryOptions& o = {},
                              void* stream = nullptr)
    {
        const nbvh_query_params p = o.params();
        check(nbvh_query(ctx_, scene_, &p, d_rays, n_rays, d_hits, stream), ctx_);
    }
    // Host buffers, synchronous.
    std::vector<nbvh_hit> neural_bvh_traversal(const std::vector<nbvh_ray>& rays, const QueryOptions& o = {})
    {
        const nbvh_query_params p = o.params();
        std::vector<nbvh_hit> hits(rays.size());
        check(nbvh_query_host(ctx_, scene_, &p, rays.data(), (uint32_t)rays.size(), hits.data()), ctx_);
        return hits;
    }
    // tcnn::Network::inference replacement: inputs n x 3P, outputs n x 8 (post-sigmoid).
    std::vector<float> network_inference(const std::vector<float>& inputs, uint32_t n, uint32_t precision,
                                         std::vector<float>* logits = nullptr)
    {
        std::vector<float> out((size_t)n * 8);
        if (logits) logits->resize((size_t)n * 8);
        check(nbvh_network_inference_host(ctx_, scene_, precision, inputs.data(), n, out.data(),
                                          logits ? logits->data() : nullptr),
              ctx_);
        return out;
    }
    // NeuralPathModule::neural_path_tracing: one 1-spp diffuse frame into d_accum (width*height
    // float4: rgb sum, sample count). `environment` (may be NULL): triangle meshes traced with the
    // neural object. Returns the rays traced per depth.
    nbvh_path_stats neural_path_tracing(const nbvh_camera& camera, uint32_t width, uint32_t height, float* d_accum,
                                        const nbvh_path_params& params, const Scene* environment = nullptr,
                                        void* stream = nullptr)
    {
        nbvh_path_stats stats{};
        check(nbvh_path_trace(ctx_, scene_, environment ? environment->get() : nullptr, &params, &camera, width,
                              height, d_accum, &stats, stream),
              ctx_);
        return stats;
    }
    // The same frame without reading the per-depth counts back (no synchronisation): with
    // params.workspace set it allocates nothing and may be captured into a HIP graph.
    void neural_path_tracing_async(const nbvh_camera& camera, uint32_t width, uint32_t height, float* d_accum,
                                   const nbvh_path_params& params, const Scene* environment = nullptr,
                                   void* stream = nullptr)
    {
        check(nbvh_path_trace(ctx_, scene_, environment ? environment->get() : nullptr, &params, &camera, width,
                              height, d_accum, nullptr, stream),
              ctx_);
    }
    // Device bytes params.workspace must hold for a frame of width x height.
    uint64_t path_trace_workspace_size(const nbvh_path_params& params, uint32_t width, uint32_t height,
                                       const Scene* environment = nullptr) const
    {
        uint64_t bytes = 0;
        check(nbvh_path_trace_workspace_size(ctx_, scene_, environment ? environment->get() : nullptr, &params, width,
                                             height, &bytes),
              ctx_);
        return bytes;
    }

private:
    nbvh_ctx* ctx_;
    nbvh_scene* scene_;
};

// Training of a scene's network (the reference's trainer->training_step): composite loss and Adam
// with the reference's defaults; commit() writes the trained network into the scene.
class Trainer {
public:
    explicit Trainer(Context& ctx, Scene& scene, uint32_t max_batch = 1u << 18, float learning_rate = 1e-2f)
        : ctx_(ctx.get())
    {
        nbvh_train_config c{};
        c.learning_rate = learning_rate;
        c.beta1         = 0.9f;
        c.beta2         = 0.999f;
        c.epsilon       = 1e-8f;
        c.l2_reg        = 1e-8f;
        c.max_batch     = max_batch;
        c.decay_start    = 10000;  // the reference's exponential learning-rate decay
        c.decay_interval = 2000;
        c.decay_base     = 0.33f;
        check(nbvh_trainer_create(ctx_, scene.get(), &c, &t_), ctx_);
    }
    ~Trainer() { nbvh_trainer_destroy(t_); }
    Trainer(const Trainer&) = delete;
    Trainer& operator=(const Trainer&) = delete;

    // One step on device buffers (n x 9 encoded segments, n x 8 targets); returns the loss.
    float training_step(const float* d_inputs, const float* d_targets, uint32_t n, void* stream = nullptr)
    {
        float loss = 0.f;
        check(nbvh_trainer_step(t_, d_inputs, d_targets, n, 0u, &loss, stream), ctx_);
        return loss;
    }
    void commit() { check(nbvh_trainer_commit(t_), ctx_); }
    nbvh_trainer* get() const { return t_; }

private:
    nbvh_ctx* ctx_ = nullptr;
    nbvh_trainer* t_ = nullptr;
};

// Tree-cut learning (the reference's training loop around the trainer): the learner grows the
// scene's N-BVH from the cut at cfg.neural_bvh_min_depth while the trainer trains the network.
class Learner {
public:
    static nbvh_learner_config defaults()
    {
        nbvh_learner_config c{};
        check(nbvh_learner_config_defaults(&c));
        return c;
    }
    Learner(Context& ctx, Scene& scene, const nbvh_learner_config& cfg) : ctx_(ctx.get())
    {
        check(nbvh_learner_create(ctx_, scene.get(), &cfg, &l_), ctx_);
    }
    ~Learner() { nbvh_learner_destroy(l_); }
    Learner(const Learner&) = delete;
    Learner& operator=(const Learner&) = delete;

    // One step: sample the training segments, train on them, split leaves on the schedule.
    float step(Trainer& trainer, void* stream = nullptr)
    {
        float loss = 0.f;
        check(nbvh_learner_step(l_, trainer.get(), &loss, stream), ctx_);
        return loss;
    }
    nbvh_learner_state state() const
    {
        nbvh_learner_state s{};
        check(nbvh_learner_get_state(l_, &s), ctx_);
        return s;
    }

private:
    nbvh_ctx* ctx_ = nullptr;
    nbvh_learner* l_ = nullptr;
};

// LDR-FLIP mean error of a rendering against a reference rendering (device float4 accumulation
// buffers); ppd 0: the reference's viewing geometry.
inline float flip(Context& ctx, const float* d_reference, const float* d_test, uint32_t width, uint32_t height,
                  float* d_error_map = nullptr, float ppd = 0.f, void* stream = nullptr)
{
    float mean = 0.f;
    check(nbvh_flip(ctx.get(), d_reference, d_test, width, height, ppd, d_error_map, &mean, stream), ctx.get());
    return mean;
}

// Path-trace defaults of the reference's path module: max depth 5 (4 bounces), sky background.
inline nbvh_path_params path_params(uint32_t precision = NBVH_PREC_FP16, uint32_t max_depth = 5,
                                    uint32_t rr_start_depth = 0, uint32_t sample = 0)
{
    nbvh_path_params p{};
    p.query.kind          = NBVH_QUERY_NEURAL_CLOSEST;
    p.query.precision     = precision;
    p.bounce_query_flags  = NBVH_QUERY_SORT_RAYS;
    p.max_depth           = max_depth;
    p.rr_start_depth      = rr_start_depth;
    p.sample              = sample;
    p.background          = NBVH_BACKGROUND_SKY;
    p.environment_albedo[0] = p.environment_albedo[1] = p.environment_albedo[2] = 0.8f;
    return p;
}

// Continuous-batching request loop (nbvh_server_*).
class Server {
public:
    explicit Server(Context& ctx, uint32_t n_streams = 4, uint32_t max_batch_rays = 1u << 21,
                    uint64_t staging_bytes = 256ull << 20)
        : ctx_(ctx.get())
    {
        nbvh_server_config c{};
        c.n_streams      = n_streams;
        c.max_batch_rays = max_batch_rays;
        c.staging_bytes  = staging_bytes;
        check(nbvh_server_create(ctx_, &c, &s_), ctx_);
    }
    ~Server() { nbvh_server_destroy(s_); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Host buffers (staged); both must stay valid until wait() returns.
    uint64_t submit(const Scene& scene, const QueryOptions& o, const nbvh_ray* rays, uint32_t n, nbvh_hit* hits)
    {
        return submit_impl(scene, o, rays, n, hits, 0u);
    }
    // Device buffers (zero-copy).
    uint64_t submit_device(const Scene& scene, const QueryOptions& o, const nbvh_ray* d_rays, uint32_t n,
                           nbvh_hit* d_hits)
    {
        return submit_impl(scene, o, d_rays, n, d_hits, NBVH_REQUEST_DEVICE_BUFFERS);
    }
    // Blocks until the request completes (timeout_us < 0: no limit); throws Error on failure or
    // timeout (status NBVH_ERR_TIMEOUT leaves the ticket valid).
    nbvh_request_info wait(uint64_t ticket, int64_t timeout_us = -1)
    {
        nbvh_request_info info{};
        check(nbvh_request_wait(s_, ticket, timeout_us, &info), ctx_);
        return info;
    }
    nbvh_server_stats stats()
    {
        nbvh_server_stats st{};
        check(nbvh_server_get_stats(s_, &st), ctx_);
        return st;
    }

private:
    uint64_t submit_impl(const Scene& scene, const QueryOptions& o, const nbvh_ray* rays, uint32_t n, nbvh_hit* hits,
                         uint32_t flags)
    {
        const nbvh_query_params p = o.params();
        uint64_t t = 0;
        check(nbvh_request_submit(s_, scene.get(), &p, rays, n, hits, flags, &t), ctx_);
        return t;
    }
    nbvh_ctx* ctx_;
    nbvh_server* s_ = nullptr;
};

}  // namespace nbvh

#endif /* NBVH_HPP_ */
