// nbvh_internal.h -- plain structures shared by the host library (g++) and the HIP kernels (hipcc).
#pragma once

#include <stdint.h>
#include "../../include/nbvh.h"

namespace nbvh {

constexpr int      kMaxLevels        = 32;
constexpr int      kNeuralStackSize  = 24;   // NEURAL_STACK_SIZE, include/cuda/neural/accel/neural_bvh.cuh:22
constexpr int      kTriStackSize     = 64;   // reference: 32 without overflow check (bvh2.cuh:25)
constexpr float    kFixedNodeDilation  = 1e-5f;  // neural_bvh_renderer_constants.cuh:30
constexpr float    kNodeExtentDilation = 1e-4f;  // neural_bvh_renderer_constants.cuh:31
constexpr float    kInferenceMintEps   = 1e-4f;  // neural_bvh_renderer_constants.cuh:25
constexpr float    kInferenceMaxtEps   = 1e-3f;  // neural_bvh_renderer_constants.cuh:26
constexpr float    kDepthEpsilon       = 0.2f;   // neural_bvh_module.h:270
constexpr float    kFixedDepthEpsilon  = 1e-5f;  // neural_bvh_module.h:271
constexpr uint32_t kDefaultMaxSteps    = 1000;   // m_max_traversal_steps, neural_bvh_module.h:200

// One level of the multiresolution grid (tcnn GridEncoding, public convention).
struct GridLevel {
    float    scale;   // exp2(l*log2(b)) * N_min - 1
    uint32_t res;     // ceil(scale) + 1
    uint32_t res2;    // res * res
    uint32_t size;    // entries in the level (min(next_multiple(res^3, 8), T)); a power of two when hashed
    uint32_t offset;  // first entry of the level
    uint32_t dense;   // 1: index = x + res*y + res^2*z; 0: coherent-prime hash
};

// Everything a neural kernel needs, passed by value (no __constant__ singletons, so several
// scenes can be resident and queried concurrently).
struct NetArgs {
    uint32_t n_levels, n_features, n_points, n_inputs;  // n_inputs = P * L * F
    uint32_t n_neurons, n_hidden_layers, n_outputs, lod;
    float    seg_t[4];                 // segment point parameters t_k
    uint32_t seg_uniform, seg_seed;    // uniform distribution: t_k from the ray's LCG stream, its seed
    float    enc_center[3];            // neural scene bbox centre, (max + min) * 0.5
    float    enc_extent[3];            // bbox extent * (1 + node_extent_dilation + fixed_node_dilation)
    float    thr_logit;                // visibility threshold on the logit (0 for 0.5)
    uint32_t max_steps;                // max traversal steps (segments <= max_steps + 1)
    uint32_t image_width;              // tiled claim order of a scanline image (0: array order)
    uint32_t queue_bands;              // ray-queue bands (1..8; 0: 8), launch shape
    uint32_t blocks_per_cu;            // persistent workgroups per CU (0: as many as fit), launch shape
    const nbvh_node* nodes;            // N-BVH nodes with DILATED boxes (1 + 1e-5 about the centre)
    uint32_t n_nodes;
    uint32_t any_hit;
    const float*    tables_f32;        // [entries x F]
    const uint16_t* tables_f16;        // [entries x F] IEEE half of tables * table_scale
    uint32_t tables_f16_bytes;         // size of tables_f16 (< 2^31: buffer-resource range)
    float    table_scale;              // power of two
    float    inv_table_scale;
    const float*    mlp_f32;           // row-major matrices, concatenated
    const uint16_t* mlp_frag;          // MFMA fragment-ordered half/bf16 weights
    const uint8_t*  mlp_frag_fp8;      // FP8 e4m3 hidden-layer fragments
    float    fp8_wscale[8];            // per-layer weight dequantisation scales (FP8 variant)
    uint32_t n_dense;                  // levels [n_dense, n_levels) are all hashed (grid resolution
                                       // grows with the level, so the dense levels are a prefix)
    uint32_t depth_mode;               // nbvh_depth_mode
    uint32_t depth_linear;             // depth output read as clamp((y - 0.45) * 10, 0, 1)
    uint32_t interpolate_depth;        // closest hit: average estimates within interp_eps
    float    interp_eps;
    float    bbox_min[3], bbox_max[3]; // neural scene box (WorldT decode)
    GridLevel level[kMaxLevels];
};

struct TriArgs {
    const float*     verts;    // [nv x 3]
    const uint32_t*  indices;  // [nt x 3]
    const nbvh_node* nodes;
    const uint32_t*  prims;
    uint32_t         any_hit;
};

// One part of a multi-part launch: n rays read from `rays`, hit records written to `hits`,
// occupying claim indices [begin, begin + n) of the launch; closest or any hit per part.
struct BatchPart {
    const nbvh_ray* rays;
    nbvh_hit*       hits;
    uint32_t        begin;
    uint32_t        n;
    uint32_t        any_hit;  // the part's rays stop at the first hit (shadow rays)
    uint32_t        reserved;
};
constexpr uint32_t kMaxBatchParts = 64;

// The rays of one query launch. parts == nullptr: one array of n rays (rays -> hits).
// Otherwise n_parts (1..kMaxBatchParts) parts in device memory, sorted by begin, parts[0].begin
// == 0, laid end to end over [0, n).
struct RayBatch {
    const nbvh_ray*  rays;
    nbvh_hit*        hits;
    const BatchPart* parts;
    uint32_t         n_parts;
    uint32_t         n;
    const uint32_t*  order;  // optional claim order: claim i traces ray order[i] (a permutation of [0, n))
    const uint32_t*  live;   // optional device-side count: the batch is claims [0, min(*live, n)), n being
                             // the capacity the launch is sized for (no parts, no image order)
};

// One leaf segment of the sorted wavefront rounds (NBVH_QUERY_WAVEFRONT).
struct WfSeg {
    uint32_t ray;   // claim index of the ray in its launch
    uint32_t node;  // N-BVH leaf
    float    mint, maxt;
};

// Per-ray state of the wavefront rounds (64 B; the layout belongs to the kernels).
struct WfRayState {
    uint32_t words[16];
};

// One training step's buffers (kernels/train.hip): the network being trained (FP32 tables in
// tcnn order, row-major matrices), per-sample activation records, gradients, loss accumulator.
struct TrainArgs {
    NetArgs      net;          // shape and grid levels
    const float* tables;
    const float* mlp;
    float*       act;          // workspace: activation rows (720 x act_stride floats), then
                               // transposed matrices and per-chunk matrix-gradient partials
    uint32_t     act_stride;   // samples per activation row (max batch rounded up to 64)
    float*       grad_tables;
    float*       grad_mlp;
    float*       loss;
    float*       sample_loss;  // optional: per-sample sum of the 8 element loss values
    void*        bins;         // workspace of the binned table gradient (train_bin_bytes)
};

// Tree-cut learning data of one N-BVH node (NeuralTreeCutData,
// include/cuda/neural/neural_bvh_renderer.h:36-47).
struct TreeCutNode {
    float    loss;
    uint32_t iter_alive, depth, samples;
};

// Training-data sampler (granular_sample_bvh_network_input_data,
// src/cuda/neural/neural_bvh_module_kernels.cu:119-264).
struct SampleArgs {
    const nbvh_node*   nodes;          // the learner's N-BVH, boxes as built (not dilated)
    uint32_t           n_nodes;
    TriArgs            tri;            // the scene's triangles (targets)
    float              box_min[3], box_ext[3];  // neural scene bbox: origins are sampled in it
    float              enc_center[3], enc_extent[3];
    float              seg_t[4];
    uint32_t           n_points;
    float              scaling;        // sampling_bbox_scaling
    float              fixed_dilation, extent_dilation;
    float              mint_eps, maxt_eps;  // learning epsilons
    uint32_t           current_lod;    // current_learned_lod
    uint32_t           survival;       // training_node_survival
    float              survival_scale, min_survival, current_max_loss;
    const TreeCutNode* tree_cut;
    uint32_t           n_sub_batch, sub_batch_size, batch_size, step;
    float*             inputs;         // batch x (3 * n_points)
    float*             targets;        // batch x 8
    uint32_t*          node_idx;       // batch
    uint32_t*          failures;       // threads that gave up after the attempt cap
};

// LDR-FLIP (flip.hip). Accumulation buffers are float4 per pixel (rgb sum, sample count); the
// prepared buffers hold (Yy, Cx, Cz, Y / Yn) per pixel; kernels are (2r+1)^2 row-major with row =
// y offset (the y feature kernels are the transposes of the x ones).
constexpr int kFlipMaxRadius = 16;
struct FlipArgs {
    const float* ref;   // w*h float4
    const float* test;
    uint32_t     w, h;
    int          rc, rf;  // colour (contrast sensitivity) and feature kernel radii
    const float* csf;     // 3 kernels: achromatic, red-green, blue-yellow
    const float* edge;    // x kernel of the edge detector
    const float* point;   // x kernel of the point detector
    float        cmax;    // HyAB^0.7 of Hunt-adjusted green against blue
    float        rgb2xyz[9], xyz2rgb[9], white[3], inv_white[3];
    float*       prep_ref;   // w*h float4
    float*       prep_test;
    float*       map;        // optional w*h error map
    double*      partials;   // one per 16x16 tile
};

// Adam step constants (the bias correction uses each parameter's own update count).
struct AdamArgs {
    float lr, beta1, beta2, epsilon, l2_reg;
};

// One live path of the path-trace driver (32 B): its pixel, throughput and radiance so far.
struct PathState {
    uint32_t pixel;
    float    thr[3];
    float    illum[3];
    uint32_t pad;
};

// Shading parameters of the path-trace driver.
struct PathArgs {
    TriArgs  tri;                  // triangle environment (normals of triangle hits); verts NULL: none
    float*   accum;                // width*height float4
    uint32_t max_depth;
    uint32_t rr_start_depth;       // 0: no Russian roulette
    uint32_t sample;
    uint32_t background;           // nbvh_background
    float    background_color[3];
    float    env_albedo[3];
    float    normal_offset;
};

}  // namespace nbvh
