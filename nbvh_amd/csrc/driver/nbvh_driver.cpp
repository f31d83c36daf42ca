// nbvh_driver -- C++ host driver of the neural ray query through include/nbvh.hpp.
//
// The reference drives its query from C++ (NeuralBVHRenderer -> NeuralPathModule); this is the
// same path for a C++ caller of libnbvh_hip.so:
//   nbvh_driver host                      host-side builders only (runs without a GPU)
//   nbvh_driver query [options]           build the 1024x1024 icosphere workload, trace it, write
//                                         the hit records (48 B each) to --out
// options: --width W --height H --precision fp32|fp16|bf16 --kind closest|any|tri|tri-any
//          --leaves K --seed S --out FILE --repeat R (prints ms per query)
//   nbvh_driver path [options]            path-traced frames of the icosphere workload
//   nbvh_driver learn [--steps N --batch B --width W --height H --spp S]
//                                         tree-cut learning + training from the N-BVH root, then
//                                         FLIP of the neural against the triangle rendering
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nbvh.hpp"

namespace {

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

int host_selftest()
{
    nbvh::Mesh m = nbvh::icosphere(3, 7, 0.1f);
    std::vector<nbvh_node> nodes(2 * m.n_tris());
    std::vector<uint32_t> prims(m.n_tris());
    uint32_t nn = 0;
    nbvh::check(nbvh_bvh_build(m.verts.data(), m.n_verts(), m.tris.data(), m.n_tris(), 2, nodes.data(),
                               (uint32_t)nodes.size(), &nn, prims.data()));
    std::vector<nbvh_node> cut(2 * 64);
    uint32_t nc = 0;
    nbvh::check(nbvh_nbvh_cut(nodes.data(), nn, 64, cut.data(), (uint32_t)cut.size(), &nc));
    size_t size = 0;
    nbvh::check(nbvh_nbvh_serialize(cut.data(), nc, nullptr, 0, &size));
    std::vector<char> bytes(size);
    nbvh::check(nbvh_nbvh_serialize(cut.data(), nc, bytes.data(), size, &size));
    std::vector<nbvh_node> back(nc);
    uint32_t nb = 0;
    nbvh::check(nbvh_nbvh_parse(bytes.data(), size, back.data(), nc, &nb));
    if (nb != nc || std::memcmp(back.data(), cut.data(), nc * sizeof(nbvh_node)) != 0) {
        std::fprintf(stderr, "round trip mismatch\n");
        return 1;
    }
    uint64_t nt = 0, nm = 0;
    const nbvh_network_config cfg = nbvh::network_config();
    nbvh::check(nbvh_network_param_counts(&cfg, &nt, &nm));
    std::printf("host ok: %u triangles, %u BVH nodes, %u N-BVH nodes, %zu .nbvh bytes, %llu + %llu params\n",
                m.n_tris(), nn, nc, size, (unsigned long long)nt, (unsigned long long)nm);
    try {
        nbvh::Context ctx;
        std::printf("device: available\n");
    } catch (const nbvh::Error& e) {
        if (e.status() != NBVH_ERR_NO_DEVICE) throw;
        std::printf("device: none (%s)\n", e.what());
    }
    return 0;
}

int query(int argc, char** argv)
{
    uint32_t W = 1024, H = 1024, leaves = 256, repeat = 0;
    uint64_t seed = 1;
    std::string prec = "fp16", kind = "closest", out;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string a = argv[i], v = argv[i + 1];
        if (a == "--width") W = (uint32_t)std::stoul(v);
        else if (a == "--height") H = (uint32_t)std::stoul(v);
        else if (a == "--precision") prec = v;
        else if (a == "--kind") kind = v;
        else if (a == "--leaves") leaves = (uint32_t)std::stoul(v);
        else if (a == "--seed") seed = std::stoull(v);
        else if (a == "--out") out = v;
        else if (a == "--repeat") repeat = (uint32_t)std::stoul(v);
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    nbvh::QueryOptions o;
    o.precision   = prec == "fp32" ? NBVH_PREC_FP32 : prec == "bf16" ? NBVH_PREC_BF16 : NBVH_PREC_FP16;
    o.kind        = kind == "any" ? NBVH_QUERY_NEURAL_ANY
                  : kind == "tri" ? NBVH_QUERY_TRIANGLE_CLOSEST
                  : kind == "tri-any" ? NBVH_QUERY_TRIANGLE_ANY : NBVH_QUERY_NEURAL_CLOSEST;
    o.image_width = W;

    nbvh::Context ctx;
    nbvh::Mesh m = nbvh::icosphere(6, 7, 0.1f);
    nbvh::Scene scene(ctx, m);
    scene.build_neural_bvh(leaves);
    scene.init_network(nbvh::network_config(), seed, 1e-4f);
    nbvh::NeuralBVHModule module(ctx, scene);

    const float eye[3] = {0, 0, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    const nbvh_camera cam = nbvh::camera_look_at(eye, target, up, 45.f, W, H);
    const uint32_t n = W * H;
    nbvh_ray* d_rays = nullptr;
    nbvh_hit* d_hits = nullptr;
    hip_ok(hipMalloc((void**)&d_rays, (size_t)n * sizeof(nbvh_ray)), "hipMalloc rays");
    hip_ok(hipMalloc((void**)&d_hits, (size_t)n * sizeof(nbvh_hit)), "hipMalloc hits");
    hipStream_t stream;
    hip_ok(hipStreamCreate(&stream), "hipStreamCreate");
    ctx.raygen(cam, W, H, d_rays, 0, 0, stream);
    module.neural_bvh_traversal(d_rays, n, d_hits, o, stream);
    hip_ok(hipStreamSynchronize(stream), "sync");
    if (repeat) {
        hipEvent_t e0, e1;
        hip_ok(hipEventCreate(&e0), "event");
        hip_ok(hipEventCreate(&e1), "event");
        hip_ok(hipEventRecord(e0, stream), "record");
        for (uint32_t r = 0; r < repeat; r++) module.neural_bvh_traversal(d_rays, n, d_hits, o, stream);
        hip_ok(hipEventRecord(e1, stream), "record");
        hip_ok(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        std::printf("%s %s: %.4f ms per query, %.1f Mrays/s\n", prec.c_str(), kind.c_str(), ms / repeat,
                    n / (ms / repeat * 1e-3) / 1e6);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    std::vector<nbvh_hit> hits(n);
    hip_ok(hipMemcpy(hits.data(), d_hits, (size_t)n * sizeof(nbvh_hit), hipMemcpyDeviceToHost), "D2H");
    uint64_t neural = 0;
    for (const auto& h : hits) neural += NBVH_HIT_KIND(h.info) != NBVH_HIT_NONE;
    std::printf("%u rays, %llu hits\n", n, (unsigned long long)neural);
    if (!out.empty()) {
        std::ofstream f(out, std::ios::binary);
        f.write(reinterpret_cast<const char*>(hits.data()), (std::streamsize)(hits.size() * sizeof(nbvh_hit)));
    }
    (void)hipStreamDestroy(stream);
    (void)hipFree(d_rays);
    (void)hipFree(d_hits);
    return 0;
}

// Path-traced frames of the benchmark icosphere (nbvh_path_trace); writes the averaged image as
// a binary PPM when --out is given.
int path(int argc, char** argv)
{
    uint32_t W = 640, H = 480, spp = 4, depth = 5, rr = 3;
    std::string prec = "fp16", out;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string a = argv[i], v = argv[i + 1];
        if (a == "--width") W = (uint32_t)std::stoul(v);
        else if (a == "--height") H = (uint32_t)std::stoul(v);
        else if (a == "--precision") prec = v;
        else if (a == "--spp") spp = (uint32_t)std::stoul(v);
        else if (a == "--max-depth") depth = (uint32_t)std::stoul(v);
        else if (a == "--rr-start") rr = (uint32_t)std::stoul(v);
        else if (a == "--out") out = v;
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    nbvh::Context ctx;
    nbvh::Mesh m = nbvh::icosphere(6, 7, 0.1f);
    nbvh::Scene scene(ctx, m);
    scene.build_neural_bvh(256);
    scene.init_network(nbvh::network_config(), 1, 1e-4f);
    nbvh::NeuralBVHModule module(ctx, scene);
    const float eye[3] = {0, 0, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    const nbvh_camera cam = nbvh::camera_look_at(eye, target, up, 45.f, W, H);
    const uint32_t prec_id = prec == "fp32" ? NBVH_PREC_FP32 : prec == "bf16" ? NBVH_PREC_BF16 : NBVH_PREC_FP16;
    float* d_accum = nullptr;
    hip_ok(hipMalloc((void**)&d_accum, (size_t)W * H * 4 * sizeof(float)), "hipMalloc accum");
    uint64_t rays = 0;
    for (uint32_t s = 0; s < spp; s++) {
        const nbvh_path_stats st = module.neural_path_tracing(cam, W, H, d_accum, nbvh::path_params(prec_id, depth, rr, s));
        for (uint32_t d = 0; d < depth; d++) rays += st.rays[d];
    }
    std::vector<float> img((size_t)W * H * 4);
    hip_ok(hipMemcpy(img.data(), d_accum, img.size() * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    double sum = 0.0;
    for (size_t i = 0; i < (size_t)W * H; i++) sum += (img[4 * i] + img[4 * i + 1] + img[4 * i + 2]) / (3.0 * img[4 * i + 3]);
    std::printf("%ux%u, %u spp, %s: %llu rays traced, mean radiance %.5f\n", W, H, spp, prec.c_str(),
                (unsigned long long)rays, sum / ((double)W * H));
    if (!out.empty()) {
        std::ofstream f(out, std::ios::binary);
        f << "P6\n" << W << " " << H << "\n255\n";
        for (size_t i = 0; i < (size_t)W * H; i++)
            for (int c = 0; c < 3; c++) {
                const float v = img[4 * i + c] / img[4 * i + 3];
                const float g = v <= 0.f ? 0.f : v >= 1.f ? 1.f : std::pow(v, 1.f / 2.2f);
                f.put((char)(unsigned char)(g * 255.f + 0.5f));
            }
    }
    (void)hipFree(d_accum);
    return 0;
}

// The reference's way of producing a network, from C++: tree-cut learning of the icosphere's
// N-BVH from its root with the training step, then the FLIP of a neural rendering against the
// triangle rendering of the object (the validation mode's comparison).
int learn(int argc, char** argv)
{
    uint32_t steps = 1000, batch = 1u << 16, W = 320, H = 240, spp = 4;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string a = argv[i], v = argv[i + 1];
        if (a == "--steps") steps = (uint32_t)std::stoul(v);
        else if (a == "--batch") batch = (uint32_t)std::stoul(v);
        else if (a == "--width") W = (uint32_t)std::stoul(v);
        else if (a == "--height") H = (uint32_t)std::stoul(v);
        else if (a == "--spp") spp = (uint32_t)std::stoul(v);
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    nbvh::Context ctx;
    nbvh::Mesh m = nbvh::icosphere(6, 7, 0.1f);
    nbvh::Scene scene(ctx, m);
    scene.init_network(nbvh::network_config(8, 4, 18, 8, 2.f), 1, 1e-4f);  // the shipped chess network shape
    nbvh_learner_config cfg = nbvh::Learner::defaults();
    cfg.batch_size           = batch;
    cfg.max_training_steps   = steps;
    cfg.last_split_iteration = (int32_t)(steps * 3 / 4);
    cfg.splitting_interval   = 25;
    nbvh::Trainer trainer(ctx, scene, batch);
    nbvh::Learner learner(ctx, scene, cfg);
    const auto t0 = std::chrono::steady_clock::now();
    float loss = 0.f;
    for (uint32_t s = 0; s < steps; s++) loss = learner.step(trainer);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    trainer.commit();
    const nbvh_learner_state st = learner.state();
    std::printf("learned %u steps in %.1f ms (%.3f ms/step): loss %.5f, N-BVH %u nodes (%u splits), "
                "%u sampler failures\n", steps, ms, ms / steps, loss, st.n_nodes, st.total_splits, st.sample_failures);
    nbvh::NeuralBVHModule module(ctx, scene);
    const float eye[3] = {0, 0, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    const nbvh_camera cam = nbvh::camera_look_at(eye, target, up, 45.f, W, H);
    float *d_neural = nullptr, *d_tri = nullptr;
    hip_ok(hipMalloc((void**)&d_neural, (size_t)W * H * 16), "hipMalloc accum");
    hip_ok(hipMalloc((void**)&d_tri, (size_t)W * H * 16), "hipMalloc accum");
    for (uint32_t s = 0; s < spp; s++) {
        nbvh_path_params p = nbvh::path_params(NBVH_PREC_FP16, 3, 2, s);
        module.neural_path_tracing(cam, W, H, d_neural, p);
        p.query.kind = NBVH_QUERY_TRIANGLE_CLOSEST;  // the object's own triangles (ignore_neural_blas)
        for (int c = 0; c < 3; c++) p.environment_albedo[c] = 0.9f;
        module.neural_path_tracing(cam, W, H, d_tri, p);
    }
    const float err = nbvh::flip(ctx, d_tri, d_neural, W, H);
    std::printf("FLIP of the neural rendering against the triangle rendering (%ux%u, %u spp): %.4f\n", W, H, spp, err);
    (void)hipFree(d_neural);
    (void)hipFree(d_tri);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: nbvh_driver host | query [options] | path [options] | learn [options]\n");
        return 2;
    }
    try {
        if (std::strcmp(argv[1], "host") == 0) return host_selftest();
        if (std::strcmp(argv[1], "query") == 0) return query(argc, argv);
        if (std::strcmp(argv[1], "path") == 0) return path(argc, argv);
        if (std::strcmp(argv[1], "learn") == 0) return learn(argc, argv);
    } catch (const nbvh::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "unknown command %s\n", argv[1]);
    return 2;
}
