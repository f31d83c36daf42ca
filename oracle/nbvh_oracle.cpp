// nbvh_oracle.cpp -- CPU restatement of the reference's neural ray query, used ONLY as the checker.
//
// This file is test infrastructure. Only the tests, the smoke check and the benchmark's CPU
// baseline leg load it (via oracle/liboracle.so); the product library never links or calls it.
//
// It restates, function by function, the path of the reference (paths relative to its root):
//   LCG / murmur3 seeding        include/cuda/utils/lcg_rng.cuh:12-72
//   camera_raygen + jitter       include/cuda/neural/neural_bvh_renderer_kernels.cuh:74-79,
//                                include/cuda/neural/neural_bvh_module_common.cuh:230-236
//   slab test (Majercik 2018)    include/cuda/base/ray.cuh:116-135, include/base/accel/bvh.h:120-154
//   node dilation                include/cuda/neural/accel/neural_bvh2.cuh:69-117
//   NeuralBVH2::intersect        include/cuda/neural/accel/neural_bvh2.cuh:170-292
//   segment encoding             src/cuda/neural/neural_bvh_module_kernels.cu:10-97,266-303
//   hash grid (tcnn HashGrid)    config src/cuda/neural/neural_bvh_module.cu:170-191 (tcnn absent: the
//                                public tiny-cuda-nn GridEncoding convention is restated, see header notes)
//   FullyFusedMLP                src/cuda/neural/neural_bvh_module.cu:232-238 (no bias, ReLU, Sigmoid out)
//   output decode                include/cuda/neural/neural_bvh_module_common.cuh:23-57,160-209,
//                                src/cuda/neural/path/path_inference_kernels.cu:61-131
//   per-ray step loop            include/cuda/neural/neural_bvh_module_common.cuh:497-550,
//                                src/cuda/neural/path/path_module.cu:54-137
//   diffuse bounce, path shading src/cuda/neural/path/path_inference_kernels.cu:224-429 (diffuse
//                                materials), include/cuda/base/material.cuh:101-120,193-201,
//                                include/cuda/utils/rendering.cuh:11-40
//   triangle BVH closest/any     include/cuda/accel/bvh2.cuh:13-88 (with the watertight test of
//                                Woop, Benthin, Wald, JCGT 2013 in place of Moller-Trumbore,
//                                include/cuda/accel/bvh.cuh:98-136, as the build specifies)
//
// Precision modes: traversal is always FP32 (the reference's arithmetic, bit-for-bit); the
// encoding + hash grid + MLP + decode run in float (mode 0, the op sequence the GPU FP32 path
// uses, fmaf where it uses FMA) or in double (mode 1, the FP64 oracle).
//
// Pinning (see tests/test_oracle_*.py): murmur3 against published MurmurHash3_x86_32 vectors;
// LCG against its closed form; slab test against brute-force/interval arithmetic; traversal
// against brute-force leaf enumeration and the shipped chess .nbvh; grid against textbook
// trilinear interpolation on dense levels and an independent numpy hash; MLP against numpy FP64
// GEMMs; triangle BVH against brute force over all triangles; exp against numpy.
// Parity unpinned (no reference implementation or fixture exists): tcnn's exact grid
// convention and RepeatedComposite feature order, and the depth decode's double-precision
// literal (the reference evaluates u*(len+1e-3)+mint in double; the FP64 mode does the same,
// the FP32 mode uses 1e-3f).

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr uint32_t NO_NEURAL_CHILD   = 0xFFFFFFFFu;
constexpr int      NEURAL_STACK_SIZE = 24;   // include/cuda/neural/accel/neural_bvh.cuh:22
constexpr int      TRI_STACK_SIZE    = 64;
constexpr float    FIXED_NODE_DILATION    = 1e-5f;  // neural_bvh_renderer_constants.cuh:30
constexpr float    NODE_EXTENT_DILATION   = 1e-4f;  // neural_bvh_renderer_constants.cuh:31
constexpr float    INFERENCE_MINT_EPSILON = 1e-4f;  // :25
constexpr float    INFERENCE_MAXT_EPSILON = 1e-3f;  // :26
constexpr float    DEPTH_EPSILON          = 0.2f;   // neural_bvh_module.h:270
constexpr float    FIXED_DEPTH_EPSILON    = 1e-5f;  // neural_bvh_module.h:271

struct Node {
    float    mn[3];
    uint32_t prim_count;
    float    mx[3];
    uint32_t left;
};
static_assert(sizeof(Node) == 32, "node layout");

struct Ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
};
static_assert(sizeof(Ray) == 32, "ray layout");

struct Hit {
    float    t;
    uint32_t prim_id;
    uint32_t node_id;
    uint32_t info;
    float    normal_or_uv[3];
    float    logit;
    float    albedo[3];
    float    reserved;
};
static_assert(sizeof(Hit) == 48, "hit layout");

// glm::min / glm::max semantics (NaN propagation is asymmetric and is restated exactly).
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }

// ------------------------------------------------------------------------------------------
// RNG (include/cuda/utils/lcg_rng.cuh)
// ------------------------------------------------------------------------------------------
inline uint32_t murmur_mix(uint32_t hash, uint32_t k)
{
    k *= 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    hash ^= k;
    hash = ((hash << 13) | (hash >> 19)) * 5u + 0xe6546b64u;
    return hash;
}
inline uint32_t murmur_fmix(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
inline uint32_t lcg_seed(uint32_t index, uint32_t frame, uint32_t linear)
{
    uint32_t s = murmur_mix(0u, linear);
    s          = murmur_mix(s, 512u * frame + index);
    return murmur_fmix(s);
}
inline uint32_t lcg_next(uint32_t &state)
{
    state = state * 1664525u + 1013904223u;
    return state;
}
inline float lcg_float(uint32_t &state) { return std::ldexp((float)lcg_next(state), -32); }

// ------------------------------------------------------------------------------------------
// exp / sigmoid in float: Cody-Waite reduction by ln2 (hi/lo split) + degree-7 Taylor polynomial
// in Horner form, scaled by 2^n. The GPU FP32 path evaluates the identical sequence.
// ------------------------------------------------------------------------------------------
inline float expf_cw(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.972076416015625f) return 0.f;
    const float n = std::rint(x * 1.44269502162933349609375f);
    float r = std::fma(-n, 0.693145751953125f, x);
    r       = std::fma(-n, 1.428606765330187045037746429443359375e-06f, r);
    float p = 1.98412701138295233249664306640625e-04f;  // 1/5040
    p = std::fma(p, r, 1.388888922519981861114501953125e-03f);  // 1/720
    p = std::fma(p, r, 8.3333335816860198974609375e-03f);       // 1/120
    p = std::fma(p, r, 4.1666667163372039794921875e-02f);       // 1/24
    p = std::fma(p, r, 1.666666716337203979492187500e-01f);     // 1/6
    p = std::fma(p, r, 0.5f);
    p = std::fma(p, r, 1.0f);
    p = std::fma(p, r, 1.0f);
    return std::ldexp(p, (int)n);
}
inline float sigmoid_f(float z) { return 1.f / (1.f + expf_cw(-z)); }

// sin and cos in float: Cody-Waite reduction by pi/2 with the three-part constant 1.5703125 +
// 4.838705062866211e-4 - 4.371138828673793e-8 (the first two with at most 12 significant bits,
// so q * part is exact for small quadrant counts q), then the Taylor polynomials of sin (to r^9)
// and cos (to r^10) on [-pi/4, pi/4] in Horner form and the quadrant swap. The sampler kernel
// evaluates the identical sequence.
inline void sincosf_cw(float x, float &s, float &c)
{
    const float q = std::rint(x * 0.636619772367581343f);
    float r       = std::fma(-q, 1.5703125f, x);
    r             = std::fma(-q, 4.838705062866211e-4f, r);
    r             = std::fma(-q, -4.371138828673793e-8f, r);
    const float r2 = r * r;
    float sp = std::fma(r2, 2.7557319223985893e-6f, -1.9841269841269841e-4f);  // 1/9!, -1/7!
    sp       = std::fma(sp, r2, 8.3333333333333332e-3f);                       // 1/5!
    sp       = std::fma(sp, r2, -1.6666666666666666e-1f);                      // -1/3!
    const float sr = std::fma(sp, r2 * r, r);
    float cp = std::fma(r2, -2.7557319223985890e-7f, 2.4801587301587302e-5f);  // -1/10!, 1/8!
    cp       = std::fma(cp, r2, -1.3888888888888889e-3f);                      // -1/6!
    cp       = std::fma(cp, r2, 4.1666666666666664e-2f);                       // 1/4!
    cp       = std::fma(cp, r2, -0.5f);
    const float cr = std::fma(cp, r2, 1.0f);
    const int quadrant = (int)q & 3;
    s = quadrant == 0 ? sr : (quadrant == 1 ? cr : (quadrant == 2 ? -sr : -cr));
    c = quadrant == 0 ? cr : (quadrant == 1 ? -sr : (quadrant == 2 ? -cr : sr));
}
inline double sigmoid_d(double z) { return 1.0 / (1.0 + std::exp(-z)); }

// ------------------------------------------------------------------------------------------
// Slab tests (ray.cuh:116-135, bvh.h:120-154), dilation (neural_bvh2.cuh:69-117)
// ------------------------------------------------------------------------------------------
inline void slab(const float bmin[3], const float bmax[3], const Ray &r, float &tn, float &tf)
{
    float t0[3], t1[3];
    for (int c = 0; c < 3; c++) {
        t0[c] = (bmin[c] - r.o[c]) / r.d[c];
        t1[c] = (bmax[c] - r.o[c]) / r.d[c];
    }
    float s[3], b[3];
    for (int c = 0; c < 3; c++) {
        s[c] = gmin(t0[c], t1[c]);
        b[c] = gmax(t0[c], t1[c]);
    }
    tn = gmax(gmax(s[0], s[1]), s[2]);
    tf = gmin(gmin(b[0], b[1]), b[2]);
}

inline void dilate(const Node &n, float dilation, float dmin[3], float dmax[3])
{
    for (int c = 0; c < 3; c++) {
        const float ext    = n.mx[c] - n.mn[c];
        const float center = n.mn[c] + ext * 0.5f;
        const float e2     = ext * (1.f + dilation);
        dmin[c]            = center - e2 * 0.5f;
        dmax[c]            = center + e2 * 0.5f;
    }
}

inline bool is_leaf_at_lod(const Node &n, uint32_t lod) { return n.prim_count > 0 && n.prim_count <= 2u + lod; }

// ------------------------------------------------------------------------------------------
// NeuralBVH2::intersect (neural_bvh2.cuh:170-292), local stack; by default the inference
// epsilons and the fixed dilation (the training sampler passes a sampled dilation and the
// learning epsilons).
// Returns 1 (leaf segment found), 0 (stack exhausted), -1 (stack overflow).
// ------------------------------------------------------------------------------------------
int nbvh_intersect(const Node *nodes, uint32_t lod, const Ray &ray_in, uint32_t *stack, int &ss, float current_maxt,
                   float &seg_mint, float &seg_maxt, uint32_t &seg_node, float dilation = FIXED_NODE_DILATION,
                   float mint_eps = INFERENCE_MINT_EPSILON, float maxt_eps = INFERENCE_MAXT_EPSILON)
{
    Ray ray  = ray_in;
    ray.tmin = -FLT_MAX;
    ray.tmax = current_maxt;
    uint32_t idx = stack[--ss];
    while (true) {
        const Node &node = nodes[idx];
        if (is_leaf_at_lod(node, lod) || node.left == NO_NEURAL_CHILD) {
            float dmin[3], dmax[3], mint, maxt;
            dilate(node, dilation, dmin, dmax);
            slab(dmin, dmax, ray, mint, maxt);
            if (mint <= maxt) {
                bool valid = mint < ray.tmax;
                if (mint < 0.f)
                    valid = valid && (maxt > maxt_eps);
                else
                    valid = valid && (mint > mint_eps);
                if (valid) {
                    seg_mint = mint;
                    seg_maxt = maxt;
                    seg_node = idx;
                    return 1;
                }
            }
            if (ss == 0) return 0;
            idx = stack[--ss];
            continue;
        }
        uint32_t c1 = node.left, c2 = node.left + 1;
        float    m1, m2;
        {
            float dmin[3], dmax[3], tn, tf;
            dilate(nodes[c1], dilation, dmin, dmax);
            slab(dmin, dmax, ray, tn, tf);
            float mint = gmax(ray.tmin, tn), maxt = gmin(ray.tmax, tf);
            m1         = (mint <= maxt) ? mint : FLT_MAX;
            dilate(nodes[c2], dilation, dmin, dmax);
            slab(dmin, dmax, ray, tn, tf);
            mint = gmax(ray.tmin, tn);
            maxt = gmin(ray.tmax, tf);
            m2   = (mint <= maxt) ? mint : FLT_MAX;
        }
        if (m1 > m2) {
            std::swap(m1, m2);
            std::swap(c1, c2);
        }
        if (m1 == FLT_MAX) {
            if (ss == 0) return 0;
            idx = stack[--ss];
        } else {
            idx = c1;
            if (m2 != FLT_MAX) {
                stack[ss++] = c2;
                if (ss >= NEURAL_STACK_SIZE) return -1;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Network description
// ------------------------------------------------------------------------------------------
}  // namespace

extern "C" {
typedef struct oracle_net {
    uint32_t     n_levels, n_features, log2_hashmap_size, base_resolution;
    float        per_level_scale;
    uint32_t     n_points, distribution, n_neurons, n_hidden_layers, n_outputs;
    const float *tables;
    const float *mlp;
    float        bbox_min[3], bbox_max[3];
    // Decoding (pt_target_const, include/cuda/neural/neural_bvh_module.h:243-276): depth mode
    // (0 LocalT, 1 WorldT), linear depth activation, depth interpolation and its epsilon.
    uint32_t     depth_mode, depth_linear, interpolate_depth;
    float        interp_eps;
    // Uniform segment distribution (distribution 2): ray i's LCG stream is lcg_seed(segment_seed,
    // 0, i) (neural_bvh_module_kernels.cu:53-67, the query's stream seeding).
    uint32_t     segment_seed;
} oracle_net;
}

namespace {

struct Level {
    float    scale;
    uint32_t res;
    uint32_t size;    // entries (params_in_level)
    uint32_t offset;  // entries
    bool     dense;
};

inline uint32_t next_multiple(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// tiny-cuda-nn GridEncoding (public): grid_scale / grid_resolution / params_in_level / offsets.
std::vector<Level> make_levels(const oracle_net &net)
{
    std::vector<Level> lv(net.n_levels);
    const float log2s  = std::log2(net.per_level_scale);
    const uint32_t T   = 1u << net.log2_hashmap_size;
    uint32_t offset    = 0;
    for (uint32_t l = 0; l < net.n_levels; l++) {
        const float scale   = std::exp2((float)l * log2s) * (float)net.base_resolution - 1.0f;
        const uint32_t res  = (uint32_t)std::ceil(scale) + 1u;
        const uint32_t maxp = 0xFFFFFFFFu / 2u;
        uint32_t params     = std::pow((float)res, 3.f) > (float)maxp ? maxp : res * res * res;
        params              = next_multiple(params, 8u);
        params              = std::min(params, T);
        lv[l].scale         = scale;
        lv[l].res           = res;
        lv[l].size          = params;
        lv[l].offset        = offset;
        // grid_index: dense iff the stride loop finishes without exceeding the level size.
        uint64_t stride = 1;
        for (int d = 0; d < 3 && stride <= params; d++) stride *= res;
        lv[l].dense = !(params < stride);
        offset += params;
    }
    return lv;
}

inline uint32_t grid_index(const Level &L, const uint32_t g[3])
{
    uint32_t stride = 1, index = 0;
    for (int d = 0; d < 3 && stride <= L.size; d++) {
        index += g[d] * stride;
        stride *= L.res;
    }
    if (!L.dense) index = (g[0] * 1u) ^ (g[1] * 2654435761u) ^ (g[2] * 805459861u);
    return index % L.size;
}

template <typename R> inline R fma_r(R a, R b, R c);
template <> inline float fma_r<float>(float a, float b, float c) { return std::fma(a, b, c); }
template <> inline double fma_r<double>(double a, double b, double c) { return a * b + c; }

// Segment encoding (neural_bvh_module_kernels.cu:10-97,266-303). mode: float or double.
template <typename R>
void encode_segment(const oracle_net &net, const Ray &ray, float mint, float maxt, R *pts, uint32_t *rng)
{
    if (mint < 0.f) std::swap(mint, maxt);
    R eo[3], ei[3];
    const float dil = (1.f + NODE_EXTENT_DILATION) + FIXED_NODE_DILATION;
    for (int c = 0; c < 3; c++) {
        const float center = (net.bbox_max[c] + net.bbox_min[c]) * 0.5f;
        const float ext    = (net.bbox_max[c] - net.bbox_min[c]) * dil;
        const R po = (R)ray.o[c] + (R)mint * (R)ray.d[c];
        const R pi = (R)ray.o[c] + (R)maxt * (R)ray.d[c];
        eo[c]      = (po - (R)center) / (R)ext + (R)0.5;
        ei[c]      = (pi - (R)center) / (R)ext + (R)0.5;
    }
    const uint32_t P = net.n_points;
    float delta, offset;
    if (net.distribution == 1) {
        delta  = (P == 1) ? 1.f : 1.f / (float)(P - 1);
        offset = 0.f;
    } else {
        delta  = 1.f / (float)P;
        offset = 0.5f;
    }
    for (uint32_t k = 0; k < P; k++) {
        // Uniform: t = lcg_randomf(rng) per point (neural_bvh_module_kernels.cu:53-67)
        const float t = net.distribution == 2 ? lcg_float(*rng) : delta * ((float)k + offset);
        for (int c = 0; c < 3; c++) pts[3 * k + c] = ((R)1 - (R)t) * eo[c] + (R)t * ei[c];
    }
}

// Hash grid for one point: out[l*F + f].
template <typename R>
void grid_point(const oracle_net &net, const std::vector<Level> &lv, const R x[3], R *out)
{
    const uint32_t F = net.n_features;
    for (uint32_t l = 0; l < net.n_levels; l++) {
        const Level &L = lv[l];
        R        frac[3];
        uint32_t g[3];
        for (int d = 0; d < 3; d++) {
            const R pos = fma_r<R>((R)L.scale, x[d], (R)0.5);
            const R fl  = std::floor(pos);
            g[d]        = (uint32_t)(int)fl;
            frac[d]     = pos - fl;
        }
        R acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t c = 0; c < 8; c++) {
            R        w = 1;
            uint32_t gl[3];
            for (int d = 0; d < 3; d++) {
                if (c & (1u << d)) {
                    w *= frac[d];
                    gl[d] = g[d] + 1;
                } else {
                    w *= (R)1 - frac[d];
                    gl[d] = g[d];
                }
            }
            const uint32_t idx = grid_index(L, gl);
            const float   *v   = net.tables + ((size_t)L.offset + idx) * F;
            for (uint32_t f = 0; f < F; f++) acc[f] = fma_r<R>(w, (R)v[f], acc[f]);
        }
        for (uint32_t f = 0; f < F; f++) out[l * F + f] = acc[f];
    }
}

// Largest network the oracle evaluates: P * L * F inputs, hidden width, outputs.
constexpr uint32_t kMaxInputs = 4 * 32 * 8, kMaxNeurons = 256;

// One dense layer y[j] = sum_k W[j * in + k] * a[k] (row-major W), each output's sum formed in k
// order by its own chain of multiply-adds (the kernels' order); 8 outputs at a time, so that
// their independent chains overlap. relu: max(y, 0).
template <typename R>
void dense_layer(const float *W, const R *a, uint32_t in, uint32_t out, R *y, bool relu)
{
    uint32_t j = 0;
    for (; j + 8 <= out; j += 8) {
        R acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = 0; k < in; k++) {
            const R ak = a[k];
            for (uint32_t u = 0; u < 8; u++) acc[u] = fma_r<R>((R)W[(j + u) * in + k], ak, acc[u]);
        }
        for (uint32_t u = 0; u < 8; u++) y[j + u] = relu ? (acc[u] > (R)0 ? acc[u] : (R)0) : acc[u];
    }
    for (; j < out; j++) {
        R acc = 0;
        for (uint32_t k = 0; k < in; k++) acc = fma_r<R>((R)W[j * in + k], a[k], acc);
        y[j] = relu ? (acc > (R)0 ? acc : (R)0) : acc;
    }
}

// MLP: logits z[n_outputs] from features x[P*L*F].
template <typename R>
void mlp(const oracle_net &net, const R *x, R *z)
{
    const uint32_t K0 = net.n_points * net.n_levels * net.n_features;
    const uint32_t H  = net.n_neurons;
    if (K0 > kMaxInputs || H > kMaxNeurons) std::abort();
    R a[kMaxInputs], b[kMaxNeurons];
    for (uint32_t k = 0; k < K0; k++) a[k] = x[k];
    const float *W  = net.mlp;
    uint32_t     in = K0;
    for (uint32_t layer = 0; layer < net.n_hidden_layers; layer++) {
        dense_layer<R>(W, a, in, H, b, true);
        W += (size_t)H * in;
        in = H;
        for (uint32_t k = 0; k < H; k++) a[k] = b[k];
    }
    dense_layer<R>(W, a, in, net.n_outputs, z, false);
}

template <typename R>
void network_eval(const oracle_net &net, const std::vector<Level> &lv, const R *pts, R *z)
{
    const uint32_t LF = net.n_levels * net.n_features;
    if (net.n_points * LF > kMaxInputs) std::abort();
    R feat[kMaxInputs];
    for (uint32_t k = 0; k < net.n_points; k++) grid_point<R>(net, lv, pts + 3 * k, feat + k * LF);
    mlp<R>(net, feat, z);
}

inline float sig(float z) { return sigmoid_f(z); }
inline double sig(double z) { return sigmoid_d(z); }

// Depth decode, LocalT (neural_bvh_module_common.cuh:32-57). Returns FLT_MAX when rejected.
inline float decode_depth(float u, float mint, float maxt)
{
    const float len = maxt - mint;
    if (mint < 0.f) {
        const float inward = u * (len + 1e-3f);
        if (inward < maxt - len * DEPTH_EPSILON - FIXED_DEPTH_EPSILON) return maxt - inward;
        return FLT_MAX;
    }
    return u * (len + 1e-3f) + mint;
}
inline float decode_depth(double u, float mint, float maxt)
{
    // The reference writes (segment_length + 1e-3) with a double literal: evaluated in double.
    const float len = maxt - mint;
    if (mint < 0.f) {
        const float inward = (float)(u * ((double)len + 1e-3));
        if (inward < maxt - len * DEPTH_EPSILON - FIXED_DEPTH_EPSILON) return maxt - inward;
        return FLT_MAX;
    }
    return (float)(u * ((double)len + 1e-3) + (double)mint);
}

// WorldT (neural_bvh_module_common.cuh:58-84): the LocalT formulas on the ray's span [smin, smax]
// through the neural scene's box (intersect_bbox_min_max without the ray's own bounds,
// include/cuda/base/ray.cuh:116-132); `inside` is the segment's negative mint.
template <typename R>
float decode_span(R u, float smin, float smax, bool inside)
{
    const float len = smax - smin;
    if (inside) {
        const float inward = sizeof(R) == 4 ? (float)u * (len + 1e-3f) : (float)((double)u * ((double)len + 1e-3));
        if (inward < smax - len * DEPTH_EPSILON - FIXED_DEPTH_EPSILON) return smax - inward;
        return FLT_MAX;
    }
    return sizeof(R) == 4 ? (float)u * (len + 1e-3f) + smin : (float)((double)u * ((double)len + 1e-3) + (double)smin);
}

// The hit distance a segment's depth output y (post-sigmoid) decodes to under the network's
// depth mode and activation (decode_depth_estimate, neural_bvh_module_common.cuh:23-84).
template <typename R>
float decode_segment(const oracle_net &net, const Ray &ray, R y, float mint, float maxt)
{
    R u = y;
    if (net.depth_linear) {
        u = (y - (R)0.45f) * (R)10.f;
        u = u > (R)0 ? u : (R)0;
        u = u < (R)1 ? u : (R)1;
    }
    if (net.depth_mode == 1) {
        float smin, smax;
        slab(net.bbox_min, net.bbox_max, ray, smin, smax);
        return decode_span<R>(u, smin, smax, mint < 0.f);
    }
    return decode_depth(u, mint, maxt);
}

struct SegRecord {
    uint32_t ray;
    uint32_t node;
    float    mint, maxt;
    float    logit[8];
};

// One ray of the neural query (per-ray restatement of the wavefront loop: every step evaluates
// the current leaf segment, keeps the closest decoded depth and advances to the next leaf with
// current_maxt = best depth, path_module.cu:68-129, neural_bvh_module_common.cuh:497-550).
template <typename R>
void neural_ray(const oracle_net &net, const std::vector<Level> &lv, const Node *nodes, const Ray &ray, uint32_t any,
                uint32_t lod, uint32_t max_steps, float thr_logit, Hit &hit, std::vector<SegRecord> *log, uint32_t ray_idx)
{
    hit = Hit{FLT_MAX, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, {0, 0, 0}, 0.f, {0, 0, 0}, 0.f};
    float    best_t = ray.tmax;
    R        cur_n[3] = {0, 0, 0};  // the current hit's normal in the evaluation precision
    uint32_t stack[NEURAL_STACK_SIZE];
    int      ss    = 1;
    stack[0]       = 0;
    uint32_t kind = 0, flags = 0, segments = 0;
    float    mint, maxt;
    uint32_t node;
    int      st = nbvh_intersect(nodes, lod, ray, stack, ss, ray.tmax, mint, maxt, node);
    const uint32_t P = net.n_points;
    uint32_t rng = lcg_seed(net.segment_seed, 0u, ray_idx);
    while (st == 1) {
        if (segments == max_steps + 1) {
            flags |= 2u;
            break;
        }
        R pts[12], z[16];
        encode_segment<R>(net, ray, mint, maxt, pts, &rng);
        network_eval<R>(net, lv, pts, z);
        (void)P;
        segments++;
        if (log) {
            SegRecord rec{ray_idx, node, mint, maxt, {0, 0, 0, 0, 0, 0, 0, 0}};
            for (uint32_t j = 0; j < net.n_outputs && j < 8; j++) rec.logit[j] = (float)z[j];
            log->push_back(rec);
        }
        const bool occluded = z[0] > (R)thr_logit;
        if (occluded) {
            const bool inside = mint < 0.f;
            if (any) {
                bool h = true;
                if (inside) h = decode_segment<R>(net, ray, sig(z[1]), mint, maxt) != FLT_MAX;
                if (h) {
                    kind        = 1;
                    hit.node_id = node;
                    hit.logit   = (float)z[0];
                    break;
                }
            } else {
                const float depth = decode_segment<R>(net, ray, sig(z[1]), mint, maxt);
                // interpolate_depth (path_inference_kernels.cu:89-121): an estimate within the
                // epsilon of the current neural hit is averaged with it, normals summed.
                const bool near = depth != FLT_MAX && net.interpolate_depth && kind == 1 &&
                                  std::fabs(best_t - depth) <= net.interp_eps;
                if (depth != FLT_MAX && (depth < best_t || near)) {
                    const float prev_t = best_t;
                    best_t      = depth;
                    kind        = 1;
                    hit.node_id = node;
                    hit.logit   = (float)z[0];
                    R s[3], a[3];
                    for (int c = 0; c < 3; c++) {
                        s[c] = sig(z[2 + c]) * (R)2 - (R)1;
                        a[c] = sig(z[5 + c]);
                    }
                    const R inv = (R)1 / std::sqrt((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
                    R n[3];
                    for (int c = 0; c < 3; c++) {
                        n[c]          = s[c] * inv;
                        hit.albedo[c] = (float)a[c];
                    }
                    if (near) {
                        best_t = 0.5f * (prev_t + depth);
                        R m[3];
                        for (int c = 0; c < 3; c++) m[c] = n[c] + cur_n[c];
                        const R im = (R)1 / std::sqrt((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]);
                        for (int c = 0; c < 3; c++) n[c] = m[c] * im;
                    }
                    for (int c = 0; c < 3; c++) {
                        cur_n[c]            = n[c];
                        hit.normal_or_uv[c] = (float)n[c];
                    }
                }
            }
        }
        if (ss == 0) {
            st = 0;
            break;
        }
        st = nbvh_intersect(nodes, lod, ray, stack, ss, best_t, mint, maxt, node);
    }
    if (st == -1) flags |= 1u;
    hit.t    = (kind != 0 && !any) ? best_t : FLT_MAX;
    hit.info = kind | (flags << 8) | (std::min(segments, 65535u) << 16);
}

// ------------------------------------------------------------------------------------------
// Watertight ray/triangle (Woop, Benthin, Wald 2013) with FP64 recomputation of zero edge
// functions; interval [tmin, tmax) as BaseCudaBLAS::intersect_triangle (bvh.cuh:98-136).
// ------------------------------------------------------------------------------------------
struct TriSetup {
    int   kx, ky, kz;
    float Sx, Sy, Sz;
};
inline TriSetup tri_setup(const Ray &r)
{
    TriSetup s;
    int kz = 0;
    if (std::fabs(r.d[1]) > std::fabs(r.d[kz])) kz = 1;
    if (std::fabs(r.d[2]) > std::fabs(r.d[kz])) kz = 2;
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (r.d[kz] < 0.f) std::swap(kx, ky);
    s.kx = kx;
    s.ky = ky;
    s.kz = kz;
    s.Sx = r.d[kx] / r.d[kz];
    s.Sy = r.d[ky] / r.d[kz];
    s.Sz = 1.f / r.d[kz];
    return s;
}
inline bool tri_intersect(const Ray &r, const TriSetup &s, const float *v0, const float *v1, const float *v2,
                          float tmax, float &t_out, float &u_out, float &v_out)
{
    const float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    const float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    const float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    const float Ax = A[s.kx] - s.Sx * A[s.kz], Ay = A[s.ky] - s.Sy * A[s.kz];
    const float Bx = B[s.kx] - s.Sx * B[s.kz], By = B[s.ky] - s.Sy * B[s.kz];
    const float Cx = C[s.kx] - s.Sx * C[s.kz], Cy = C[s.ky] - s.Sy * C[s.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float Az = s.Sz * A[s.kz], Bz = s.Sz * B[s.kz], Cz = s.Sz * C[s.kz];
    const float T   = (U * Az + V * Bz) + W * Cz;
    const float rcp = 1.f / det;
    const float t   = T * rcp;
    if (!(t >= r.tmin && t < tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

inline float box_enter(const Node &n, const Ray &r, float tmax)
{
    float tn, tf;
    slab(n.mn, n.mx, r, tn, tf);
    const float mint = gmax(r.tmin, tn), maxt = gmin(tmax, tf);
    return (mint <= maxt) ? mint : FLT_MAX;
}

void tri_ray(const float *verts, const uint32_t *idx, const Node *nodes, const uint32_t *prims, const Ray &ray,
             uint32_t any, Hit &hit)
{
    hit = Hit{FLT_MAX, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, {0, 0, 0}, 0.f, {0, 0, 0}, 0.f};
    const TriSetup s = tri_setup(ray);
    float tmax = ray.tmax;
    uint32_t stack[TRI_STACK_SIZE];
    int ss = 0;
    uint32_t ni = 0, flags = 0;
    bool found = false;
    while (true) {
        const Node &n = nodes[ni];
        if (n.prim_count > 0) {
            for (uint32_t p = n.left; p < n.left + n.prim_count; p++) {
                const uint32_t tri = prims[p];
                float t, u, v;
                if (tri_intersect(ray, s, verts + 3 * idx[3 * tri], verts + 3 * idx[3 * tri + 1],
                                  verts + 3 * idx[3 * tri + 2], tmax, t, u, v)) {
                    found                = true;
                    tmax                 = t;
                    hit.t                = t;
                    hit.prim_id          = tri;
                    hit.normal_or_uv[0]  = u;
                    hit.normal_or_uv[1]  = v;
                    if (any) goto done;
                }
            }
            if (ss == 0) break;
            ni = stack[--ss];
            continue;
        }
        uint32_t c1 = n.left, c2 = n.left + 1;
        float m1 = box_enter(nodes[c1], ray, tmax), m2 = box_enter(nodes[c2], ray, tmax);
        if (m1 > m2) {
            std::swap(m1, m2);
            std::swap(c1, c2);
        }
        if (m1 == FLT_MAX) {
            if (ss == 0) break;
            ni = stack[--ss];
        } else {
            ni = c1;
            if (m2 != FLT_MAX) {
                if (ss == TRI_STACK_SIZE) {
                    flags |= 1u;
                    break;
                }
                stack[ss++] = c2;
            }
        }
    }
done:
    hit.info = (found ? 2u : 0u) | (flags << 8);
}

}  // namespace

// ==========================================================================================
// C entry points (ctypes)
// ==========================================================================================
ORACLE_API uint32_t oracle_murmur_mix(uint32_t h, uint32_t k) { return murmur_mix(h, k); }
ORACLE_API uint32_t oracle_murmur_fmix(uint32_t h) { return murmur_fmix(h); }
ORACLE_API uint32_t oracle_lcg_seed(uint32_t index, uint32_t frame, uint32_t linear)
{
    return lcg_seed(index, frame, linear);
}
ORACLE_API uint32_t oracle_lcg_next(uint32_t *state) { return lcg_next(*state); }
ORACLE_API float oracle_lcg_float(uint32_t *state) { return lcg_float(*state); }
ORACLE_API float oracle_expf(float x) { return expf_cw(x); }
ORACLE_API float oracle_sigmoidf(float x) { return sigmoid_f(x); }

ORACLE_API void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

ORACLE_API int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// camera: origin[3], du[3], dv[3], dir_top_left[3]; rays: n = w*h Ray records.
ORACLE_API void oracle_raygen(const float *cam, uint32_t w, uint32_t h, uint32_t sample, uint32_t frame, void *rays_out)
{
    Ray *rays = (Ray *)rays_out;
    for (uint32_t i = 0; i < w * h; i++) {
        const uint32_t px = i % w, py = i / w;
        uint32_t st        = lcg_seed(sample, frame, i);
        const float xf     = ((float)px + lcg_float(st)) / (float)w;
        const float yf     = ((float)py + lcg_float(st)) / (float)h;
        float d[3];
        for (int c = 0; c < 3; c++) d[c] = (xf * cam[3 + c] + yf * cam[6 + c]) + cam[9 + c];
        const float inv = 1.f / std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        Ray &r          = rays[i];
        for (int c = 0; c < 3; c++) {
            r.o[c] = cam[c];
            r.d[c] = d[c] * inv;
        }
        r.tmin = 1e-4f;
        r.tmax = FLT_MAX;
    }
}

// Shading normal of a hit turned toward the incoming direction d (shading_frame,
// include/cuda/utils/rendering.cuh:11-22): predicted for a neural hit, geometric for a triangle
// hit. False when the hit has no normal source.
static bool shading_normal(const float *verts, const uint32_t *tris, const Hit &h, const float d[3], float nrm[3])
{
    const uint32_t kind = h.info & 0xFFu;
    if (kind == 1u) {
        for (int c = 0; c < 3; c++) nrm[c] = h.normal_or_uv[c];
    } else if (kind == 2u && verts) {
        const float *a = verts + 3 * (size_t)tris[3 * (size_t)h.prim_id];
        const float *b = verts + 3 * (size_t)tris[3 * (size_t)h.prim_id + 1];
        const float *c = verts + 3 * (size_t)tris[3 * (size_t)h.prim_id + 2];
        const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        nrm[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nrm[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nrm[2] = e1[0] * e2[1] - e1[1] * e2[0];
    } else {
        return false;
    }
    const float len = std::sqrt((nrm[0] * nrm[0] + nrm[1] * nrm[1]) + nrm[2] * nrm[2]);
    const float sgn = ((nrm[0] * d[0] + nrm[1] * d[1]) + nrm[2] * d[2]) > 0.f ? -1.f : 1.f;
    const float inv = len > 0.f ? sgn / len : 0.f;
    for (int c = 0; c < 3; c++) nrm[c] *= inv;
    if (!(len > 0.f)) nrm[2] = 1.f;
    return true;
}

// sample_lambertian_dir(n, ortho_basis(n), u) (material.cuh:101-120,193-201; utils/common.cuh:
// 21-35): concentric disk mapping, z = sqrt(1 - r^2), into the frame of n. Returns the local z.
static float lambertian_direction(float u0, float u1, const float nrm[3], float wi[3])
{
    const float sx = 2.f * u0 - 1.f, sy = 2.f * u1 - 1.f;
    float dx = 0.f, dy = 0.f;
    if (sx != 0.f || sy != 0.f) {
        float rad, th;
        if (std::fabs(sx) > std::fabs(sy)) {
            rad = sx;
            th  = 0.785398163397448310f * (sy / sx);
        } else {
            rad = sy;
            th  = 1.570796326794896619f - 0.785398163397448310f * (sx / sy);
        }
        dx = rad * std::cos(th);
        dy = rad * std::sin(th);
    }
    float hz         = std::sqrt(std::max(0.f, 1.f - dx * dx - dy * dy));
    const float hinv = 1.f / std::sqrt((dx * dx + dy * dy) + hz * hz);
    const float hx = dx * hinv, hy = dy * hinv;
    hz *= hinv;
    float bt[3] = {0.f, 0.f, 0.f}, t[3], b[3];
    if (nrm[0] < 0.6f && nrm[0] > -0.6f)
        bt[0] = 1.f;
    else if (nrm[1] < 0.6f && nrm[1] > -0.6f)
        bt[1] = 1.f;
    else if (nrm[2] < 0.6f && nrm[2] > -0.6f)
        bt[2] = 1.f;
    else
        bt[0] = 1.f;
    t[0] = bt[1] * nrm[2] - bt[2] * nrm[1];
    t[1] = bt[2] * nrm[0] - bt[0] * nrm[2];
    t[2] = bt[0] * nrm[1] - bt[1] * nrm[0];
    float ti = 1.f / std::sqrt((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]);
    for (int c = 0; c < 3; c++) t[c] *= ti;
    b[0] = nrm[1] * t[2] - nrm[2] * t[1];
    b[1] = nrm[2] * t[0] - nrm[0] * t[2];
    b[2] = nrm[0] * t[1] - nrm[1] * t[0];
    ti = 1.f / std::sqrt((b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]);
    for (int c = 0; c < 3; c++) b[c] *= ti;
    for (int c = 0; c < 3; c++) wi[c] = (hx * t[c] + hy * b[c]) + hz * nrm[c];
    return hz;
}

// One diffuse bounce (path_inference_kernels.cu:380-393): the next ray leaves p = o + t*d + 1e-4*n
// (normal_epsilon_offset, path_module.h:220) in a Lambertian direction about the shading normal
// n, u two LCG floats seeded with (sample, bounce + 1, ray). Hit-less rays come out with tmax = -1.
ORACLE_API void oracle_scatter_diffuse(const float *verts, const uint32_t *tris, const void *rays_in,
                                       const void *hits_in, uint32_t n, uint32_t sample, uint32_t bounce,
                                       void *rays_out)
{
    const Ray *rays = (const Ray *)rays_in;
    const Hit *hits = (const Hit *)hits_in;
    Ray *out        = (Ray *)rays_out;
    for (uint32_t i = 0; i < n; i++) {
        const Ray &r = rays[i];
        Ray &q       = out[i];
        float nrm[3];
        if (!shading_normal(verts, tris, hits[i], r.d, nrm)) {
            q = Ray{{0.f, 0.f, 0.f}, 0.f, {0.f, 0.f, 1.f}, -1.f};
            continue;
        }
        uint32_t st    = lcg_seed(sample, bounce + 1u, i);
        const float u0 = lcg_float(st), u1 = lcg_float(st);
        (void)lambertian_direction(u0, u1, nrm, q.d);
        for (int c = 0; c < 3; c++) q.o[c] = (r.o[c] + hits[i].t * r.d[c]) + nrm[c] * 1e-4f;
        q.tmin = 1e-4f;
        q.tmax = FLT_MAX;
    }
}

// Path state of the path-trace restatement: pixel, throughput, radiance so far (32 B).
struct PathState {
    uint32_t pixel;
    float thr[3];
    float illum[3];
    uint32_t pad;
};

// One depth of diffuse path tracing (hybrid_scatter_and_get_next_tlas_traversal_step,
// src/cuda/neural/path/path_inference_kernels.cu:224-429, diffuse materials only):
//   no hit: illum += thr * background (background_hit_eval, :37-58; skylike or constant) and the
//     path ends; depth == max_depth: the path ends;
//   else thr *= albedo (Lambertian bsdf * |cos| / pdf), a zero pdf or albedo ends the path,
//     Russian roulette at depth >= rr_start (rr_terminate, include/cuda/utils/rendering.cuh:24-40:
//     survive with p = min(1, max(thr)), thr /= p), and the next ray leaves
//     hit + normal_offset * n in a Lambertian direction; LCG seeded (sample, depth, pixel): two
//     floats for the direction, then one for the roulette.
// An ending path adds (illum, 1) to accum[pixel] (sample 0 overwrites; accumulate_pixel_value,
// include/cuda/neural/neural_bvh_renderer_kernels.cuh:81-110). Live paths are appended in order
// to next_rays / next_state; returns their count.
ORACLE_API uint32_t oracle_path_shade(const float *verts, const uint32_t *tris, uint32_t max_depth, uint32_t rr_start,
                                      uint32_t sample, uint32_t background, const float *bg_color,
                                      const float *env_albedo, float normal_offset, uint32_t depth,
                                      const void *rays_in, const void *hits_in, const void *state_in, uint32_t n,
                                      void *next_rays_out, void *next_state_out, float *accum)
{
    const Ray *rays       = (const Ray *)rays_in;
    const Hit *hits       = (const Hit *)hits_in;
    const PathState *sts  = (const PathState *)state_in;
    Ray *next_rays        = (Ray *)next_rays_out;
    PathState *next_state = (PathState *)next_state_out;
    uint32_t m            = 0;
    auto splat = [&](const PathState &s) {
        float *px = accum + 4 * (size_t)s.pixel;
        for (int c = 0; c < 3; c++) px[c] = (sample ? px[c] : 0.f) + s.illum[c];
        px[3] = (sample ? px[3] : 0.f) + 1.f;
    };
    for (uint32_t i = 0; i < n; i++) {
        const Ray &r = rays[i];
        const Hit &h = hits[i];
        PathState s  = sts[i];
        const uint32_t kind = h.info & 0xFFu;
        if (kind == 0u) {
            float bg[3];
            if (background == 1u) {
                const float t = 0.5f * (r.d[1] + 1.f);
                bg[0] = (1.f - t) * 1.f + t * 0.5f;
                bg[1] = (1.f - t) * 1.f + t * 0.7f;
                bg[2] = (1.f - t) * 1.f + t * 1.f;
            } else {
                for (int c = 0; c < 3; c++) bg[c] = bg_color[c];
            }
            for (int c = 0; c < 3; c++) s.illum[c] = s.illum[c] + s.thr[c] * bg[c];
            splat(s);
            continue;
        }
        if (depth >= max_depth) {
            splat(s);
            continue;
        }
        float nrm[3];
        if (!shading_normal(verts, tris, h, r.d, nrm)) {
            splat(s);
            continue;
        }
        const float *alb = kind == 1u ? h.albedo : env_albedo;
        uint32_t rng     = lcg_seed(sample, depth, s.pixel);
        const float u0 = lcg_float(rng), u1 = lcg_float(rng);
        float wi[3];
        const float cos_i = lambertian_direction(u0, u1, nrm, wi);
        for (int c = 0; c < 3; c++) s.thr[c] = s.thr[c] * alb[c];
        if (!(cos_i > 0.f) || (alb[0] == 0.f && alb[1] == 0.f && alb[2] == 0.f)) {
            splat(s);
            continue;
        }
        if (rr_start != 0u && depth >= rr_start) {
            const float survive = std::min(1.f, std::max(std::max(s.thr[0], s.thr[1]), s.thr[2]));
            if (lcg_float(rng) < survive) {
                for (int c = 0; c < 3; c++) s.thr[c] = s.thr[c] / survive;
            } else {
                splat(s);
                continue;
            }
        }
        Ray &q = next_rays[m];
        for (int c = 0; c < 3; c++) {
            q.o[c] = (r.o[c] + h.t * r.d[c]) + nrm[c] * normal_offset;
            q.d[c] = wi[c];
        }
        q.tmin          = 1e-4f;
        q.tmax          = FLT_MAX;
        next_state[m++] = s;
    }
    return m;
}

// Slab test of one box (check_ray_bounds false): returns hit, writes tmin/tmax.
ORACLE_API int oracle_slab(const float *bmin, const float *bmax, const void *ray, float *tn, float *tf)
{
    slab(bmin, bmax, *(const Ray *)ray, *tn, *tf);
    return *tn <= *tf;
}

// Dilated box of one node (neural_bvh2.cuh:73-78).
ORACLE_API void oracle_dilate(const void *node, float dilation, float *dmin, float *dmax)
{
    dilate(*(const Node *)node, dilation, dmin, dmax);
}

// Leaf sequence of NeuralBVH2::intersect without culling (current_maxt = tmax of the call,
// repeated until the stack empties). Writes up to cap (node, mint, maxt); returns the count,
// or -1 on stack overflow.
ORACLE_API int oracle_nbvh_leaf_sequence(const void *nodes, uint32_t lod, const void *ray, float current_maxt,
                                         uint32_t *out_nodes, float *out_mint, float *out_maxt, uint32_t cap)
{
    const Node *N = (const Node *)nodes;
    uint32_t stack[NEURAL_STACK_SIZE];
    int ss   = 1;
    stack[0] = 0;
    uint32_t n = 0;
    while (true) {
        float mint, maxt;
        uint32_t node;
        const int st = nbvh_intersect(N, lod, *(const Ray *)ray, stack, ss, current_maxt, mint, maxt, node);
        if (st < 0) return -1;
        if (st == 0) break;
        if (n < cap) {
            out_nodes[n] = node;
            out_mint[n]  = mint;
            out_maxt[n]  = maxt;
        }
        n++;
        if (ss == 0) break;
    }
    return (int)n;
}

// Number of table entries x F, per-level scale/res/size/offset/dense (for cross-checks).
ORACLE_API uint64_t oracle_grid_levels(const oracle_net *net, float *scale, uint32_t *res, uint32_t *size,
                                       uint32_t *offset, uint32_t *dense)
{
    const auto lv = make_levels(*net);
    uint64_t total = 0;
    for (size_t l = 0; l < lv.size(); l++) {
        if (scale) scale[l] = lv[l].scale;
        if (res) res[l] = lv[l].res;
        if (size) size[l] = lv[l].size;
        if (offset) offset[l] = lv[l].offset;
        if (dense) dense[l] = lv[l].dense;
        total += lv[l].size;
    }
    return total * net->n_features;
}

// Encoded segment points for n (ray, mint, maxt): out float[n * 3P] (mode 0) / double (mode 1);
// under the uniform distribution segment i takes the first P draws of stream i.
ORACLE_API void oracle_encode(const oracle_net *net, const void *rays, const float *mint, const float *maxt, uint32_t n,
                              uint32_t mode, void *out)
{
    const Ray *R = (const Ray *)rays;
    const uint32_t D = 3 * net->n_points;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t rng = lcg_seed(net->segment_seed, 0u, i);  // uniform: segment i draws from stream i
        if (mode == 0)
            encode_segment<float>(*net, R[i], mint[i], maxt[i], (float *)out + (size_t)i * D, &rng);
        else
            encode_segment<double>(*net, R[i], mint[i], maxt[i], (double *)out + (size_t)i * D, &rng);
    }
}

// Hash-grid features for n points (x[n*3]): out[n * L*F] (float mode 0 / double mode 1).
ORACLE_API void oracle_grid(const oracle_net *net, const void *x, uint32_t n, uint32_t mode, void *out)
{
    const auto lv = make_levels(*net);
    const uint32_t LF = net->n_levels * net->n_features;
    for (uint32_t i = 0; i < n; i++) {
        if (mode == 0) {
            grid_point<float>(*net, lv, (const float *)x + 3 * i, (float *)out + (size_t)i * LF);
        } else {
            grid_point<double>(*net, lv, (const double *)x + 3 * i, (double *)out + (size_t)i * LF);
        }
    }
}

// Network on encoded inputs in[n * 3P] (float for mode 0, double for mode 1 -- both read
// FP32 inputs here): logits and sigmoid outputs, double arrays [n * n_outputs].
ORACLE_API void oracle_network(const oracle_net *net, const float *in, uint32_t n, uint32_t mode, double *logits,
                               double *outputs)
{
    const auto lv = make_levels(*net);
    const uint32_t D = 3 * net->n_points, O = net->n_outputs;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        if (mode == 0) {
            float p[12], z[16];
            for (uint32_t c = 0; c < D; c++) p[c] = in[i * D + c];
            network_eval<float>(*net, lv, p, z);
            for (uint32_t j = 0; j < O; j++) {
                if (logits) logits[i * O + j] = z[j];
                if (outputs) outputs[i * O + j] = sigmoid_f(z[j]);
            }
        } else {
            double p[12], z[16];
            for (uint32_t c = 0; c < D; c++) p[c] = in[i * D + c];
            network_eval<double>(*net, lv, p, z);
            for (uint32_t j = 0; j < O; j++) {
                if (logits) logits[i * O + j] = z[j];
                if (outputs) outputs[i * O + j] = sigmoid_d(z[j]);
            }
        }
    }
}

// Full neural query. kind: 0 closest, 1 any. mode: 0 float network, 1 double network.
// seg_log (optional, may be NULL): records {ray, node, mint, maxt, logit[8]} (48 B each) up to
// seg_cap; *n_seg receives the total count (only when n_threads == 1 is the order the ray order).
ORACLE_API void oracle_neural_query(const oracle_net *net, const void *nodes, const void *rays, uint32_t n, uint32_t kind,
                                    uint32_t mode, uint32_t lod, uint32_t max_steps, float threshold, void *hits_out,
                                    void *seg_log, uint64_t seg_cap, uint64_t *n_seg)
{
    const auto lv = make_levels(*net);
    const Node *N = (const Node *)nodes;
    const Ray *R  = (const Ray *)rays;
    Hit *H        = (Hit *)hits_out;
    const float thr_logit = (threshold == 0.5f) ? 0.f : (float)std::log((double)threshold / (1.0 - (double)threshold));
    if (seg_log || n_seg) {
        std::vector<SegRecord> log;
        for (uint32_t i = 0; i < n; i++) {
            if (mode == 0)
                neural_ray<float>(*net, lv, N, R[i], kind, lod, max_steps, thr_logit, H[i], &log, i);
            else
                neural_ray<double>(*net, lv, N, R[i], kind, lod, max_steps, thr_logit, H[i], &log, i);
        }
        if (seg_log) std::memcpy(seg_log, log.data(), std::min<uint64_t>(seg_cap, log.size()) * sizeof(SegRecord));
        if (n_seg) *n_seg = log.size();
        return;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        if (mode == 0)
            neural_ray<float>(*net, lv, N, R[i], kind, lod, max_steps, thr_logit, H[i], nullptr, (uint32_t)i);
        else
            neural_ray<double>(*net, lv, N, R[i], kind, lod, max_steps, thr_logit, H[i], nullptr, (uint32_t)i);
    }
}

// Triangle BVH query (closest: any = 0, shadow: any = 1).
ORACLE_API void oracle_tri_query(const float *verts, const uint32_t *indices, const void *nodes, const uint32_t *prims,
                                 const void *rays, uint32_t n, uint32_t any, void *hits_out)
{
    const Node *N = (const Node *)nodes;
    const Ray *R  = (const Ray *)rays;
    Hit *H        = (Hit *)hits_out;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) tri_ray(verts, indices, N, prims, R[i], any, H[i]);
}

// Brute force over all triangles (closest hit, ties to the lowest triangle id).
ORACLE_API void oracle_tri_brute(const float *verts, const uint32_t *indices, uint32_t n_tris, const void *rays,
                                 uint32_t n, void *hits_out)
{
    const Ray *R = (const Ray *)rays;
    Hit *H       = (Hit *)hits_out;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Hit h{FLT_MAX, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, {0, 0, 0}, 0.f, {0, 0, 0}, 0.f};
        const TriSetup s = tri_setup(R[i]);
        for (uint32_t tri = 0; tri < n_tris; tri++) {
            float t, u, v;
            if (tri_intersect(R[i], s, verts + 3 * indices[3 * tri], verts + 3 * indices[3 * tri + 1],
                              verts + 3 * indices[3 * tri + 2], FLT_MAX, t, u, v)) {
                if (t < h.t) {
                    h.t               = t;
                    h.prim_id         = tri;
                    h.normal_or_uv[0] = u;
                    h.normal_or_uv[1] = v;
                    h.info            = 2u;
                }
            }
        }
        H[i] = h;
    }
}

// ------------------------------------------------------------------------------------------
// Training-data sampler: granular_sample_bvh_network_input_data
// (src/cuda/neural/neural_bvh_module_kernels.cu:119-264) with
// stack_based_sample_uniform_bvh_density (include/cuda/neural/neural_bvh_module_common.cuh:703-780),
// square_to_uniform_sphere (include/cuda/utils/sampling.cuh:8-14), local_to_bbox
// (include/base/bbox.h:81-85), encode_segment_from_world_po_pi (neural_bvh_module_kernels.cu:10-97,
// no swap of a negative entry) and generate_single_bvh_network_output_reference
// (include/cuda/neural/neural_bvh_module_common.cuh:783-861), for at most one training bounce.
// The LCG draws are taken in the reference's textual order (glm's vector constructors read
// left to right). tree_cut: 4 words per node (loss, iter_alive, depth, samples).
// Outputs: inputs / targets / node_idx of the slots written (the caller passes their previous
// contents); per slot the debug record of its last sampled segment: ray, [mint, maxt], LoD,
// dilation, and whether the slot's inputs were written (1) or the sample was dropped (0).
// ------------------------------------------------------------------------------------------
extern "C" {
typedef struct oracle_sample_cfg {
    uint32_t batch_size, n_sub_batch, step, current_lod, survival;
    float    scaling, fixed_dilation, extent_dilation, mint_eps, maxt_eps;
    float    survival_scale, min_survival, current_max_loss;
    float    box_min[3], box_max[3];
    uint32_t n_points, distribution;
} oracle_sample_cfg;
}

namespace {
inline float loss_heuristic(const uint32_t *tc, uint32_t node)
{
    float loss;
    std::memcpy(&loss, tc + 4 * node, 4);
    const uint32_t iter_alive = tc[4 * node + 1], samples = tc[4 * node + 3];
    return samples == 0 ? -FLT_MAX : loss / ((float)iter_alive * std::sqrt((float)samples / (float)iter_alive));
}
}  // namespace

ORACLE_API void oracle_learn_sample(const oracle_sample_cfg *c, const void *nbvh, const uint32_t *tree_cut,
                                    const float *verts, const uint32_t *tris, const void *bvh, const uint32_t *prims,
                                    float *inputs, float *targets, uint32_t *node_idx, void *dbg_rays,
                                    float *dbg_seg, uint32_t *dbg_lod, float *dbg_dil, uint32_t *written)
{
    const Node *N   = (const Node *)nbvh;
    const Node *B   = (const Node *)bvh;
    Ray *DR         = (Ray *)dbg_rays;
    const uint32_t P = c->n_points, n_in = 3 * P;
    float seg_t[4];
    {
        float delta, offset;
        if (c->distribution == 1) {
            delta  = (P == 1) ? 1.f : 1.f / (float)(P - 1);
            offset = 0.f;
        } else {
            delta  = 1.f / (float)P;
            offset = 0.5f;
        }
        for (uint32_t k = 0; k < 4; k++) seg_t[k] = delta * ((float)k + offset);
    }
    float ext[3], center[3], enc_ext[3];
    const float enc_dil = (1.f + NODE_EXTENT_DILATION) + FIXED_NODE_DILATION;
    for (int k = 0; k < 3; k++) {
        ext[k]     = c->box_max[k] - c->box_min[k];
        center[k]  = (c->box_max[k] + c->box_min[k]) * 0.5f;
        enc_ext[k] = ext[k] * enc_dil;
    }
    const uint32_t sub_batch = c->batch_size / c->n_sub_batch;
    const uint32_t n_lods    = 7u - c->current_lod;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)sub_batch; ii++) {
        uint32_t stack[NEURAL_STACK_SIZE];
        int ss = 0;
        Ray ray{{0.f, 0.f, 0.f}, 1e-4f, {1.f, 0.f, 0.f}, FLT_MAX};
        bool cont = false;
        uint32_t slot = (uint32_t)ii * c->n_sub_batch;
        for (uint32_t inner = 0; inner < c->n_sub_batch; inner++) {
            uint32_t rng = lcg_seed(c->step * c->batch_size, 0u, slot);
            // square_to_uniform_sphere(lcg_random2f(rng))
            const float u0 = lcg_float(rng), u1 = lcg_float(rng);
            const float phi = (2.0f * 3.14159265358979323846f) * u0;
            const float ct  = u1 * 2.0f - 1.0f;
            const float st  = std::sqrt(1.0f - ct * ct);
            float sp, cp;
            sincosf_cw(phi, sp, cp);
            const float dir[3] = {st * cp, st * sp, ct};
            // sampled_lod = current + min(uint(u * n), n - 1); sampled_dilation
            const uint32_t k   = (uint32_t)(lcg_float(rng) * (float)n_lods);
            const uint32_t lod = c->current_lod + std::min(k, n_lods - 1u);
            const float dil    = c->fixed_dilation + lcg_float(rng) * c->extent_dilation;
            bool valid = cont && ss != 0;
            float mint = 0.f, maxt = 0.f;
            uint32_t node = 0;
            uint32_t attempts = 0;
            do {
                if (!valid) {
                    stack[0] = 0;
                    ss       = 1;
                    float p[3];
                    const float h = (c->scaling - 1.f) * 0.5f;
                    for (int q = 0; q < 3; q++) p[q] = c->scaling * lcg_float(rng) - h;
                    for (int q = 0; q < 3; q++) {
                        ray.o[q] = p[q] * ext[q] + c->box_min[q];
                        ray.d[q] = dir[q];
                    }
                    ray.tmin = 1e-4f;
                    ray.tmax = FLT_MAX;
                }
                valid = nbvh_intersect(N, lod, ray, stack, ss, FLT_MAX, mint, maxt, node, dil, c->mint_eps,
                                       c->maxt_eps) == 1;
                if (!valid) ++attempts;
            } while (!valid && attempts < 65536u);
            DR[slot]              = ray;
            dbg_seg[2 * slot]     = mint;
            dbg_seg[2 * slot + 1] = maxt;
            dbg_lod[slot]         = lod;
            dbg_dil[slot]         = dil;
            written[slot]         = 0;
            if (!valid) {  // the sampler gives up on this slot (never reached in practice)
                ss   = 0;
                cont = false;
                continue;
            }
            node_idx[slot] = node;
            if (c->survival) {
                const float est = loss_heuristic(tree_cut, node);
                float keep      = std::max(c->min_survival, est / c->current_max_loss);
                if (!(est <= c->current_max_loss)) keep = 1.f;
                if (c->survival_scale * keep < lcg_float(rng)) continue;
            }
            float eo[3], ei[3];
            for (int q = 0; q < 3; q++) {
                const float po = ray.o[q] + mint * ray.d[q];
                const float pi = ray.o[q] + maxt * ray.d[q];
                eo[q]          = (po - center[q]) / enc_ext[q] + 0.5f;
                ei[q]          = (pi - center[q]) / enc_ext[q] + 0.5f;
            }
            for (uint32_t pt = 0; pt < P; pt++) {
                const float t = seg_t[pt];
                for (int q = 0; q < 3; q++) inputs[(size_t)slot * n_in + 3 * pt + q] = (1.f - t) * eo[q] + t * ei[q];
            }
            Ray lr  = ray;
            lr.tmin = mint;
            lr.tmax = maxt;
            Hit h;
            tri_ray(verts, tris, B, prims, lr, 0u, h);
            float *out     = targets + (size_t)slot * 8;
            const bool hit = (h.info & 0xFFu) == 2u;
            for (int j = 0; j < 8; j++) out[j] = 0.f;
            if (hit) {
                out[0] = 1.f;
                out[1] = (float)((double)(h.t - mint) / ((double)(maxt - mint) + 1e-3));
                const float *a = verts + 3 * tris[3 * h.prim_id], *b = verts + 3 * tris[3 * h.prim_id + 1],
                            *d = verts + 3 * tris[3 * h.prim_id + 2];
                const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const float e2[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
                const float nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                const float inv = 1.f / std::sqrt((nr[0] * nr[0] + nr[1] * nr[1]) + nr[2] * nr[2]);
                for (int q = 0; q < 3; q++) {
                    out[2 + q] = (nr[q] * inv + 1.f) * 0.5f;
                    out[5 + q] = 0.9f;
                }
            }
            written[slot] = 1;
            cont          = !hit;
            slot++;
        }
    }
}

ORACLE_API void oracle_sincosf(const float *x, uint32_t n, float *s, float *c)
{
    for (uint32_t i = 0; i < n; i++) sincosf_cw(x[i], s[i], c[i]);
}
