/*
 * nbvh.h -- C ABI of the MI355X-native N-BVH neural ray query (libnbvh_hip.so).
 *
 * The library replaces the inference side of the reference's Neural BVH Module:
 *   load an N-BVH and network weights, submit a ray batch, get hit records.
 * Every entry point returns an nbvh_status (0 = NBVH_OK); no exception crosses the ABI and
 * nothing calls exit(). The message of the last failure on a context (or the last failure of a
 * context-free call on this thread) is returned by nbvh_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   class NeuralBVHModule                      include/cuda/neural/neural_bvh_module.h:144-241
 *   NeuralPathModule::neural_bvh_traversal      src/cuda/neural/path/path_module.cu:54-137
 *   NeuralBVHRenderer::load_neural_bvh          src/cuda/neural/neural_bvh_renderer_accel.cpp:208-254
 *   NeuralBVHRenderer::save_neural_bvh          src/cuda/neural/neural_bvh_renderer_accel.cpp:170-206
 *   NeuralBVHRenderer::load_network_model       src/cuda/neural/neural_bvh_renderer_init.cu:394-404
 *   tcnn::Network::inference (9 x n -> 8 x n)   call site src/cuda/neural/path/path_module.cu:76-92
 *   CudaBVH2::intersect<ShadowRay>              include/cuda/accel/bvh2.cuh:13-88
 *   CudaBackend::build_blas (BVH2, leaf <= 2)   src/cuda/cuda_backend.cpp:186-310
 *   traverse_and_build_neural_bvh               src/cuda/neural/neural_bvh_renderer_accel.cpp:25-88
 *   camera_raygen + get_lcg_rng jitter          include/cuda/neural/neural_bvh_renderer_kernels.cuh:74-79,
 *                                               include/cuda/neural/neural_bvh_module_common.cuh:230-236
 *   CudaBackend::update_camera                  src/cuda/cuda_backend.cpp:595-611
 */
#ifndef NBVH_H_
#define NBVH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBVH_ABI_VERSION 2u

typedef enum nbvh_status {
    NBVH_OK                  = 0,
    NBVH_ERR_INVALID_ARGUMENT = 1, /* null pointer, size out of range, bad enum */
    NBVH_ERR_HIP             = 2,  /* a HIP runtime call failed */
    NBVH_ERR_NO_DEVICE       = 3,  /* no usable GPU */
    NBVH_ERR_OUT_OF_MEMORY   = 4,  /* device allocation failed or HBM budget exceeded */
    NBVH_ERR_FORMAT          = 5,  /* corrupt or truncated .nbvh / parameter blob */
    NBVH_ERR_STATE           = 6,  /* e.g. neural query on a scene without N-BVH or weights */
    NBVH_ERR_UNSUPPORTED     = 7,  /* configuration the kernels do not implement */
    NBVH_ERR_TIMEOUT         = 8,  /* nbvh_request_wait timed out */
    NBVH_ERR_CAPACITY        = 9   /* caller buffer too small; required size returned */
} nbvh_status;

/* ---------------------------------------------------------------------------------------------
 * Plain data types (identical layouts on host and device).
 * ------------------------------------------------------------------------------------------- */

/* 32-byte BVH node, identical to the reference's BVH2Node / NeuralBVH2Node
 * (include/base/accel/bvh.h:66-73, include/cuda/neural/accel/neural_bvh2.cuh:33).
 * Interior: prim_count == 0, children at left_or_prim_offset and left_or_prim_offset + 1.
 * Triangle-BVH leaf: prim_count primitives starting at left_or_prim_offset in the index list.
 * N-BVH leaf: prim_count in [1, 8] (1 = LEAF_FINAL, 2 + l = LEAF_LOD(l)) or
 *             left_or_prim_offset == NBVH_NO_NEURAL_CHILD. */
typedef struct nbvh_node {
    float    aabb_min[3];
    uint32_t prim_count;
    float    aabb_max[3];
    uint32_t left_or_prim_offset;
} nbvh_node;

#define NBVH_NO_NEURAL_CHILD 0xFFFFFFFFu
#define NBVH_LEAF_FINAL      1u
#define NBVH_LEAF_LOD(l)     (2u + (l))
#define NBVH_NUM_LODS        7u

/* Ray, identical to the reference's Ray3f (include/cuda/base/ray.cuh:65-79). */
typedef struct nbvh_ray {
    float o[3];
    float tmin; /* triangle queries: closest/any hit in [tmin, tmax); neural queries ignore it */
    float d[3];
    float tmax; /* neural queries: leaves entered at or beyond tmax are skipped, depths must be < tmax */
} nbvh_ray;

enum nbvh_hit_kind { NBVH_HIT_NONE = 0, NBVH_HIT_NEURAL = 1, NBVH_HIT_TRIANGLE = 2 };

enum nbvh_hit_flag {
    NBVH_FLAG_STACK_OVERFLOW = 1u, /* traversal stack exceeded; the ray stopped (reference: silent miss) */
    NBVH_FLAG_STEP_CAP       = 2u  /* max_traversal_steps reached; the ray stopped */
};

/* 48-byte hit record: the reference's NeuralTraversalResult (path_module.h:19-42) merged with
 * PreliminaryItsData (ray.cuh:22-33), the way HybridIntersectionResult (path_module.h:46-129)
 * overlays them (normal_or_uv, albedo_or_mesh_prim_id). */
typedef struct nbvh_hit {
    float    t;          /* hit distance, FLT_MAX when nothing was hit */
    uint32_t prim_id;    /* TRIANGLE: triangle index; else 0xFFFFFFFF */
    uint32_t node_id;    /* NEURAL: N-BVH leaf of the accepted segment; else 0xFFFFFFFF */
    uint32_t info;       /* kind (bits 0-7) | flags (bits 8-15) | leaf segments evaluated (16-31, saturating) */
    float    normal_or_uv[3];           /* NEURAL: normalize(2*out[2:5]-1); TRIANGLE: (u, v, 0) */
    float    logit;                     /* NEURAL: visibility logit of the accepted segment */
    float    albedo[3];                 /* NEURAL: out[5:8]; TRIANGLE: 0 */
    float    reserved;
} nbvh_hit;

#define NBVH_HIT_KIND(info)     ((info) & 0xFFu)
#define NBVH_HIT_FLAGS(info)    (((info) >> 8) & 0xFFu)
#define NBVH_HIT_SEGMENTS(info) ((info) >> 16)

/* Multiresolution hash grid (tcnn "HashGrid", CoherentPrime hash, linear interpolation),
 * shared by every segment point (RepeatedComposite, Concatenation):
 * src/cuda/neural/neural_bvh_module.cu:170-191, include/cuda/neural/utils/tcnn_config.h:69-110. */
typedef struct nbvh_grid_config {
    uint32_t n_levels;             /* L, <= 32 */
    uint32_t n_features_per_level; /* F in {1, 2, 4, 8} */
    uint32_t log2_hashmap_size;    /* log2 T, <= 26 */
    uint32_t base_resolution;      /* N_min */
    float    per_level_scale;      /* b */
} nbvh_grid_config;

enum nbvh_segment_distribution {
    NBVH_SEGMENT_EQUIDISTANT_INTERIOR = 0, /* t_k = (k + 0.5) / P */
    NBVH_SEGMENT_EQUIDISTANT_EXTERIOR = 1, /* t_k = k / (P - 1) (every shipped config) */
    /* t_k drawn for every evaluated segment from the ray's LCG stream (lcg_randomf), the stream
     * seeded as get_lcg_rng(segment_seed, 0, i) for ray i of its array or part (Uniform,
     * src/cuda/neural/neural_bvh_module_kernels.cu:53-67). Queries only: path tracing, wavefront
     * rounds and the learner refuse it (NBVH_ERR_UNSUPPORTED). */
    NBVH_SEGMENT_UNIFORM = 2
};

/* Network: segment -> P points -> hash grid -> MLP (no biases, ReLU hidden, Sigmoid output).
 * include/cuda/neural/neural_bvh_module.h:79-131 (segment encoding) and
 * include/cuda/neural/utils/tcnn_config.h:228-264 (FullyFusedMLP). */
/* How the depth output decodes into a hit distance (DepthLearningMode,
 * include/cuda/neural/neural_bvh_module_common.cuh:23-150; the reference's Position and
 * LocalPosition modes need a 3-wide depth output and are not provided). */
enum nbvh_depth_mode {
    NBVH_DEPTH_LOCAL_T = 0, /* along the segment: mint + u * (length + 1e-3) (the reference's default) */
    NBVH_DEPTH_WORLD_T = 1  /* along the ray's span through the neural scene's box instead of the segment */
};

typedef struct nbvh_network_config {
    nbvh_grid_config grid;
    uint32_t n_points;             /* P segment points, 1..4 (reference: 3) */
    uint32_t segment_distribution; /* nbvh_segment_distribution */
    uint32_t n_neurons;            /* hidden width, 64 */
    uint32_t n_hidden_layers;      /* 4 (=> 5 matrices) */
    uint32_t n_outputs;            /* 8: vis, depth, normal3, albedo3 */
    /* Decoding (pt_target_const, include/cuda/neural/neural_bvh_module.h:243-276); all zero is the
     * reference's default and the only form the trainer takes. */
    uint32_t depth_mode;              /* nbvh_depth_mode */
    uint32_t depth_linear_activation; /* output_depth_linear_activation: u = clamp((y - 0.45) * 10, 0, 1) */
    uint32_t interpolate_depth;       /* interpolate_depth (closest hit): an estimate within
                                       * depth_interpolation_eps of the current neural hit is averaged
                                       * with it, normals summed and renormalised
                                       * (src/cuda/neural/path/path_inference_kernels.cu:89-121) */
    float    depth_interpolation_eps; /* 0 => 1e-3 */
} nbvh_network_config;

/* Arithmetic of the neural query / network inference. */
enum nbvh_precision {
    NBVH_PREC_FP32 = 0, /* SIMT FP32, FP32 tables: the plain path */
    NBVH_PREC_FP16 = 1, /* MFMA FP16 in / FP32 accumulate, FP16 tables */
    NBVH_PREC_BF16 = 2, /* MFMA BF16 in / FP32 accumulate, FP16 tables */
    NBVH_PREC_FP8  = 3  /* MFMA FP16 first/last layer, FP8 e4m3 hidden layers with per-layer scales */
};

enum nbvh_query_kind {
    NBVH_QUERY_NEURAL_CLOSEST   = 0, /* N-BVH + network, closest decoded depth */
    NBVH_QUERY_NEURAL_ANY       = 1, /* N-BVH + network, first occluded segment */
    NBVH_QUERY_TRIANGLE_CLOSEST = 2, /* triangle BVH fallback, closest hit */
    NBVH_QUERY_TRIANGLE_ANY     = 3  /* triangle BVH fallback, any hit (shadow rays) */
};

typedef struct nbvh_query_params {
    uint32_t kind;                /* nbvh_query_kind */
    uint32_t precision;           /* nbvh_precision (neural kinds) */
    uint32_t lod;                 /* rendered_lod, 0..6 */
    uint32_t max_traversal_steps; /* reference m_max_traversal_steps; 0 => 1000 */
    float    visibility_threshold;/* primary_visibility_threshold on the sigmoid output; 0 => 0.5 */
    uint32_t image_width;         /* scheduling hint: the rays are an image of this width in scanline
                                     order (width % 8 == 0, n_rays % (8*width) == 0); they are then
                                     traced in 8x8 pixel tiles. 0 => array order. Results do not
                                     depend on it. */
    uint32_t flags;               /* nbvh_query_flags */
    uint32_t segment_seed;        /* seed of the rays' LCG streams under NBVH_SEGMENT_UNIFORM
                                     (ignored otherwise) */
} nbvh_query_params;

enum nbvh_query_flags {
    /* Trace the rays in a coherence order: sorted on the device by direction octant and the Morton
     * code of the origin (per part of a multi-part launch) before the launch. Hits stay at each
     * ray's own index and are identical; worthwhile for incoherent batches (bounces, shadow rays).
     * Uses stream-ordered transient device memory (hipMallocAsync) of about 24 B per ray, so a
     * sorted query is not meant for HIP graph capture. A scene whose launch shape has ray_order 1
     * (nbvh_autotune sets it where sorting was faster) gets it without the flag for batches of at
     * least 65,536 rays that carry no image_width hint (nbvh_query_noalloc sorts only when asked). */
    NBVH_QUERY_SORT_RAYS = 1u,
    /* Neural kinds: evaluate in sorted wavefront rounds instead of the persistent kernel, the
     * step-loop schedule of NeuralPathModule::neural_bvh_traversal
     * (src/cuda/neural/path/path_module.cu:54-137). Each round traverses every active ray to its
     * next leaf (one thread per ray), sorts the round's segments by the Morton code of their
     * midpoints and evaluates them in that order. Same hit records as the persistent kernel. On
     * 4M incoherent rays it is slower than NBVH_QUERY_SORT_RAYS (about 29 ms vs 21 ms: one round
     * per segment of the longest ray, each with a host synchronisation); it is kept for callers
     * that want the reference's round structure. Synchronises the stream once per round (two
     * 4-byte reads) and uses stream-ordered transient memory of about 300 B per ray; not for
     * graph capture. */
    NBVH_QUERY_WAVEFRONT = 2u,
    /* Trace the rays in array order, whatever the scene's launch shape says (ray_order). */
    NBVH_QUERY_ARRAY_ORDER = 4u
};

/* Pinhole camera in the reference's CameraParams form (origin, du, dv, dir_top_left). */
typedef struct nbvh_camera {
    float origin[3];
    float du[3];
    float dv[3];
    float dir_top_left[3];
} nbvh_camera;

typedef struct nbvh_ctx   nbvh_ctx;
typedef struct nbvh_scene nbvh_scene;

typedef struct nbvh_create_info {
    int32_t  device;          /* HIP device ordinal */
    uint32_t reserved0;
    uint64_t hbm_budget_bytes;/* 0 => 90% of free device memory at creation */
} nbvh_create_info;

/* ---------------------------------------------------------------------------------------------
 * Library / context.
 * ------------------------------------------------------------------------------------------- */
uint32_t    nbvh_abi_version(void);
const char* nbvh_status_string(int32_t status);
/* Last failure message of ctx, or of the calling thread's last context-free call if ctx is NULL. */
const char* nbvh_last_error(const nbvh_ctx* ctx);

int32_t nbvh_create(const nbvh_create_info* info, nbvh_ctx** out_ctx);
int32_t nbvh_destroy(nbvh_ctx* ctx);
/* HBM accounting of the context (bytes). */
int32_t nbvh_memory_usage(const nbvh_ctx* ctx, uint64_t* used_bytes, uint64_t* budget_bytes);

/* ---------------------------------------------------------------------------------------------
 * Host-side builders (no GPU needed).
 * ------------------------------------------------------------------------------------------- */

/* Procedural icosphere: 20*4^s triangles, unit radius scaled by `radius`, vertices displaced
 * radially by amplitude * (3-octave value noise in [-1, 1], seed), centred at center[3].
 * With verts/indices NULL only the counts are returned. */
int32_t nbvh_mesh_icosphere(uint32_t subdivisions, uint64_t seed, float amplitude, const float center[3],
                            float radius, float* verts, uint32_t* n_verts, uint32_t* indices, uint32_t* n_tris);

/* Geodesic sphere of frequency k: each icosahedron face split into a k x k triangle grid,
 * 20*k^2 triangles, 10*k^2+2 vertices, displaced like nbvh_mesh_icosphere. */
int32_t nbvh_mesh_geosphere(uint32_t frequency, uint64_t seed, float amplitude, const float center[3],
                            float radius, float* verts, uint32_t* n_verts, uint32_t* indices, uint32_t* n_tris);

/* Binned-SAH BVH2 over triangles (max_leaf_size primitives per leaf; reference uses 2,
 * include/cuda/cuda_backend.h:232). nodes needs 2*n_tris-1 entries, prim_indices n_tris. */
int32_t nbvh_bvh_build(const float* verts, uint32_t n_verts, const uint32_t* indices, uint32_t n_tris,
                       uint32_t max_leaf_size, nbvh_node* nodes, uint32_t node_capacity, uint32_t* n_nodes,
                       uint32_t* prim_indices);

/* Cut a BVH2 to an N-BVH of at most n_leaves leaves (largest surface area split first).
 * Cut nodes become LEAF_LOD(0) leaves, BVH leaves LEAF_FINAL (reference flags,
 * neural_bvh_renderer_accel.cpp:47-52). out needs 2*n_leaves-1 entries. */
int32_t nbvh_nbvh_cut(const nbvh_node* bvh_nodes, uint32_t n_bvh_nodes, uint32_t n_leaves, nbvh_node* out,
                      uint32_t capacity, uint32_t* n_out);

/* .nbvh (bvh-v2 serialisation of Node<float,3,64,32>): parse / write.
 * nbvh_nbvh_parse with out == NULL returns the node count only. */
int32_t nbvh_nbvh_parse(const void* bytes, size_t size, nbvh_node* out, uint32_t capacity, uint32_t* n_nodes);
int32_t nbvh_nbvh_serialize(const nbvh_node* nodes, uint32_t n_nodes, void* out, size_t capacity, size_t* size);

/* Number of FP32 parameters of the hash tables and of the MLP for a network configuration. */
int32_t nbvh_network_param_counts(const nbvh_network_config* cfg, uint64_t* n_table_params, uint64_t* n_mlp_params);
/* Seeded initialisation: tables ~ U(+-table_range), MLP Kaiming-uniform U(+-sqrt(6/fan_in)).
 * MLP layout: W1 [n_neurons x P*L*F], W2..W_H [n_neurons x n_neurons], W_out [n_outputs x n_neurons],
 * each row-major (out x in), concatenated. */
int32_t nbvh_network_init_params(const nbvh_network_config* cfg, uint64_t seed, float table_range, float* tables,
                                 float* mlp);

/* Weights in tiny-cuda-nn's flat parameter order, as the reference's trainer serialises them
 * (NeuralBVHRenderer::save_network_model / load_network_model,
 * src/cuda/neural/neural_bvh_renderer_init.cu:385-404: the "params_binary" float array of
 * Trainer::serialize). Assumed order (public tiny-cuda-nn; no reference weights file or test
 * exists to confirm it): NetworkWithInputEncoding puts the FullyFusedMLP's matrices first, each
 * row-major [out x in] with the input width padded to a multiple of 16 and the output matrix padded
 * to a multiple of 16 rows, then the grid's tables (level-major, entry-major, the order of
 * nbvh_scene_load_network's tables). to_network drops the padding; to_tcnn writes it as zeros.
 * n_params must equal the count for cfg (NBVH_ERR_INVALID_ARGUMENT otherwise). */
int32_t nbvh_tcnn_params_to_network(const nbvh_network_config* cfg, const float* params, uint64_t n_params,
                                    float* tables, float* mlp);
/* params NULL: only *n_params is written. */
int32_t nbvh_network_to_tcnn_params(const nbvh_network_config* cfg, const float* tables, const float* mlp,
                                    float* params, uint64_t capacity, uint64_t* n_params);

/* Pinhole camera from eye/target/up, vertical fov (degrees) and image size (update_camera). */
int32_t nbvh_camera_look_at(const float eye[3], const float target[3], const float up[3], float fovy_degrees,
                            uint32_t width, uint32_t height, nbvh_camera* out);

/* ---------------------------------------------------------------------------------------------
 * Device-resident scenes.
 * ------------------------------------------------------------------------------------------- */

/* Upload a triangle mesh and build its BVH2 (leaf <= 2). The neural scene bbox defaults to the
 * BVH root box. */
int32_t nbvh_scene_create(nbvh_ctx* ctx, const float* verts, uint32_t n_verts, const uint32_t* indices,
                          uint32_t n_tris, nbvh_scene** out_scene);
/* Scene without triangles (an N-BVH + weights only, e.g. loaded from a .nbvh). */
int32_t nbvh_scene_create_empty(nbvh_ctx* ctx, nbvh_scene** out_scene);
int32_t nbvh_scene_destroy(nbvh_scene* scene);

/* Copy out the triangle BVH (nodes, prim_indices) of the scene. */
int32_t nbvh_scene_get_bvh(const nbvh_scene* scene, nbvh_node* nodes, uint32_t node_capacity, uint32_t* n_nodes,
                           uint32_t* prim_indices, uint32_t prim_capacity, uint32_t* n_prims);

/* N-BVH: cut from the triangle BVH, parse a .nbvh blob, or set nodes directly. */
int32_t nbvh_scene_build_nbvh(nbvh_scene* scene, uint32_t n_leaves);
int32_t nbvh_scene_load_nbvh(nbvh_scene* scene, const void* bytes, size_t size);
int32_t nbvh_scene_set_nbvh(nbvh_scene* scene, const nbvh_node* nodes, uint32_t n_nodes);
int32_t nbvh_scene_get_nbvh(const nbvh_scene* scene, nbvh_node* nodes, uint32_t capacity, uint32_t* n_nodes);
/* Encoding domain (neural_scene_bbox, neural_bvh_renderer_accel.cpp:553-554): min[3], max[3]. */
int32_t nbvh_scene_set_bbox(nbvh_scene* scene, const float bbox_min[3], const float bbox_max[3]);
int32_t nbvh_scene_get_bbox(const nbvh_scene* scene, float bbox_min[3], float bbox_max[3]);

/* Network weights (FP32 master copy; FP16/BF16/FP8 forms are derived on upload). */
int32_t nbvh_scene_load_network(nbvh_scene* scene, const nbvh_network_config* cfg, const float* tables,
                                uint64_t n_table_params, const float* mlp, uint64_t n_mlp_params);
int32_t nbvh_scene_init_network(nbvh_scene* scene, const nbvh_network_config* cfg, uint64_t seed, float table_range);
int32_t nbvh_scene_get_network(const nbvh_scene* scene, nbvh_network_config* cfg, float* tables,
                               uint64_t table_capacity, float* mlp, uint64_t mlp_capacity);

/* ---------------------------------------------------------------------------------------------
 * Queries (device pointers, asynchronous on `stream`, a hipStream_t; NULL = default stream).
 * ------------------------------------------------------------------------------------------- */

/* Primary rays: pixel i = y*width + x, jitter from the reference's LCG seeded by
 * (index = sample, frame = frame, linear = i). */
int32_t nbvh_raygen(nbvh_ctx* ctx, const nbvh_camera* camera, uint32_t width, uint32_t height, uint32_t sample,
                    uint32_t frame, nbvh_ray* d_rays, void* stream);

/* One diffuse bounce of a path: for every hit, the next ray from the hit point (offset 1e-4 along
 * the shading normal) in a cosine-distributed direction about it (the reference's Lambertian
 * sampler, material.cuh:101-120,193-201, path_inference_kernels.cu:380-393). Shading normal:
 * the predicted normal of a neural hit, the geometric normal of a triangle hit, turned toward
 * the incoming ray. Random numbers: the raygen LCG with frame = bounce + 1. Rays without a hit
 * come out inactive with tmax = -1 (compact them before the next query). Device pointers. */
int32_t nbvh_scatter_diffuse(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_ray* d_rays, const nbvh_hit* d_hits,
                             uint32_t n, uint32_t sample, uint32_t bounce, nbvh_ray* d_next, void* stream);

/* Ground-truth network targets of n segments (ray i clipped to [d_mint[i], d_maxt[i])) from the
 * scene's triangles, as the reference labels training data
 * (generate_single_bvh_network_output_reference, neural_bvh_module_common.cuh:783-861, LocalT
 * depth :586-629): d_targets[i*8 + j] = visibility (1/0), depth (t - mint) / ((maxt - mint) +
 * 1e-3), normal (n + 1) / 2 (geometric normal), albedo (default base colour 0.9); zeros when the
 * segment holds no triangle. Compare with the network's sigmoid outputs for label-vs-neural
 * diagnostics. Device pointers, stream-ordered. */
int32_t nbvh_segment_labels(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_ray* d_rays, const float* d_mint,
                            const float* d_maxt, uint32_t n, float* d_targets, void* stream);

/* One query over n rays already in device memory; hit records to d_hits. */
int32_t nbvh_query(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_query_params* params, const nbvh_ray* d_rays,
                   uint32_t n_rays, nbvh_hit* d_hits, void* stream);

/* nbvh_query without any allocation. No counterpart in the reference: this entry point is for
 * callers that cannot allocate device memory (code under HIP graph capture, allocation-free
 * serving loops). Same arguments and hit records as nbvh_query; it calls no hipMalloc, hipFree or
 * allocating library routine. Its scratch (the queue counters and, with NBVH_QUERY_SORT_RAYS, the
 * sort keys and claim order) is one static 24 MiB device array per GPU, set up by nbvh_create and
 * held until the process ends. Limits: a sorted query takes at most NBVH_NOALLOC_MAX_SORTED_RAYS
 * rays (NBVH_ERR_CAPACITY beyond); NBVH_QUERY_WAVEFRONT is refused (NBVH_ERR_UNSUPPORTED). Every
 * call on one GPU shares the array, whatever its context: calls that overlap on one GPU (other
 * streams, other threads, other contexts) race and must be ordered by the caller. */
#define NBVH_NOALLOC_MAX_SORTED_RAYS 524288u
int32_t nbvh_query_noalloc(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_query_params* params,
                           const nbvh_ray* d_rays, uint32_t n_rays, nbvh_hit* d_hits, void* stream);

/* One launch over several device ray arrays (parts), each closest-hit or any-hit: e.g. the
 * closest-hit and shadow rays of a wavefront bounce in one persistent kernel. params->kind picks
 * the family (neural or triangle) and its precision / lod / limits; each part's any_hit picks
 * closest (0) or any (1) hit. Stream-ordered like nbvh_query; the part table is staged in a ring
 * of 2048 per-context slots, so at most 2048 such calls may be in flight per context. */
typedef struct nbvh_query_part {
    const nbvh_ray* rays;    /* device pointer */
    nbvh_hit*       hits;    /* device pointer */
    uint32_t        n_rays;
    uint32_t        any_hit; /* 0: closest hit, 1: any hit (shadow ray) */
} nbvh_query_part;

int32_t nbvh_query_parts(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_query_params* params,
                         const nbvh_query_part* parts, uint32_t n_parts, void* stream);

/* Launch shape of the persistent query kernels: how many workgroups per CU stay resident, into
 * how many bands the ray queue is split (each XCD starts on its own band, so neighbouring rays
 * share an L2) and the order in which batches without an image-order hint are traced. Results
 * never depend on it. Kept per scene and precision. */
typedef struct nbvh_launch_shape {
    uint32_t blocks_per_cu; /* 0: as many as fit */
    uint32_t queue_bands;   /* 1..8; 0: 8 */
    uint32_t ray_order;     /* 0: array order; 1: batches of at least 65,536 rays with image_width 0 and
                             * neither NBVH_QUERY_SORT_RAYS nor NBVH_QUERY_ARRAY_ORDER are traced in
                             * the coherence order of NBVH_QUERY_SORT_RAYS (not while the stream is
                             * captured) */
    uint32_t reserved;
} nbvh_launch_shape;

int32_t nbvh_scene_set_launch_shape(nbvh_scene* scene, uint32_t precision, const nbvh_launch_shape* shape);
int32_t nbvh_scene_get_launch_shape(const nbvh_scene* scene, uint32_t precision, nbvh_launch_shape* shape);

/* Autotuner: times the candidate launch shapes (workgroups per CU in {max, 1, 2, 3} x bands in
 * {1, 2, 4, 8}) of `params` on the given device rays, then, for a batch the coherence order
 * applies to (see ray_order), the fastest shape again with the rays sorted (the sort timed
 * too), and keeps the fastest for the scene and params->precision. Synchronous on `stream`;
 * d_hits receives the last run's (shape-independent) hits. best / best_ms may be NULL. */
int32_t nbvh_autotune(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_query_params* params, const nbvh_ray* d_rays,
                      uint32_t n_rays, nbvh_hit* d_hits, void* stream, nbvh_launch_shape* best, float* best_ms);
/* The same over the parts of an nbvh_query_parts launch (the batch the caller will run, e.g. a
 * bounce's closest-hit and shadow rays together). No counterpart in the reference, whose launch
 * shapes are fixed at compile time. */
int32_t nbvh_autotune_parts(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_query_params* params,
                            const nbvh_query_part* parts, uint32_t n_parts, void* stream, nbvh_launch_shape* best,
                            float* best_ms);

/* Hybrid query over one neural object and a scene of ordinary triangle meshes (the reference's
 * hybrid TLAS of the neural BLAS and mesh BLASes, tlas_traversal_and_fill_next_hybrid_traversal,
 * include/cuda/neural/neural_bvh_module_common.cuh:291-420): closest hit over both, or for
 * NBVH_QUERY_NEURAL_ANY occlusion by either. The triangles are traced first; the neural query then
 * looks only in front of each ray's triangle hit, and a neural hit replaces the triangle record.
 * params->kind is NBVH_QUERY_NEURAL_CLOSEST or _ANY. Stream-ordered; uses transient
 * stream-ordered memory (hipMallocAsync) of 80 B per ray (not meant for HIP graph capture). */
int32_t nbvh_query_hybrid(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* triangles,
                          const nbvh_query_params* params, const nbvh_ray* d_rays, uint32_t n_rays, nbvh_hit* d_hits,
                          void* stream);

/* ---------------------------------------------------------------------------------------------
 * Path tracing: the reference's NeuralPathModule::neural_path_tracing
 * (src/cuda/neural/path/path_module.cu:139-245) with its scatter step
 * hybrid_scatter_and_get_next_tlas_traversal_step (src/cuda/neural/path/path_inference_kernels.cu:
 * 224-429), restricted to diffuse materials and a procedural background.
 * ------------------------------------------------------------------------------------------- */

/* What an escaped ray sees (background_hit_eval, path_inference_kernels.cu:37-58). */
enum nbvh_background {
    NBVH_BACKGROUND_CONSTANT = 0, /* background_color (black = the reference's envmap_eval off) */
    NBVH_BACKGROUND_SKY      = 1  /* skylike_background: a = (d.y + 1)/2, (1-a)*(1,1,1) + a*(0.5,0.7,1) */
};

typedef struct nbvh_path_params {
    nbvh_query_params query;      /* query of the primary rays. kind NBVH_QUERY_NEURAL_CLOSEST: the
                                   * object through its N-BVH and network; NBVH_QUERY_TRIANGLE_CLOSEST:
                                   * the object's own triangles instead, diffuse with
                                   * environment_albedo (the reference's ignore_neural_blas, the
                                   * triangle rendering its validation compares with; no environment) */
    uint32_t bounce_query_flags;  /* nbvh_query_flags of the bounce rays (query.flags is for depth 1) */
    uint32_t max_depth;           /* path vertices, 1..16; primary hits have depth 1 (m_max_depth) */
    uint32_t rr_start_depth;      /* Russian roulette at depths >= this (rr_terminate,
                                   * include/cuda/utils/rendering.cuh:24-40); 0: off */
    uint32_t sample;              /* 0: the frame overwrites d_accum; > 0: it adds to it */
    uint32_t background;          /* nbvh_background */
    float background_color[3];
    float environment_albedo[3];  /* diffuse albedo of the triangle environment's meshes */
    float normal_offset;          /* next origin = hit + offset * shading normal; 0: 1e-4
                                   * (normal_epsilon_offset, include/cuda/neural/path/path_module.h:220) */
    uint32_t path_lod_switch;     /* nonzero: rays at depth > 1 use N-BVH LoD lod_switch instead of
                                   * query.lod (path_lod_switch / lod_switch,
                                   * include/cuda/neural/neural_bvh_module_common.cuh:333-335) */
    uint32_t lod_switch;
    float secondary_visibility_threshold; /* visibility threshold of the rays at depth > 1
                                   * (secondary_visibility_threshold, neural_bvh_module.h:268-269, picked
                                   * after the first vertex, neural_bvh_module_common.cuh:171-173);
                                   * 0: query.visibility_threshold at every depth */
    uint32_t reserved;
    void*    workspace;           /* device memory of at least nbvh_path_trace_workspace_size bytes
                                   * for the frame's buffers; NULL: transient stream-ordered memory
                                   * (not while the stream is captured) */
    uint64_t workspace_bytes;
} nbvh_path_params;

typedef struct nbvh_path_stats {
    uint32_t rays[16];     /* rays traced at depth d in rays[d - 1] */
    uint32_t reserved[16];
} nbvh_path_stats;

/* One frame of 1-spp diffuse path tracing into d_accum (width*height float4, pixel y*width + x:
 * rgb radiance sum and the sample count in w, like accumulate_pixel_value,
 * include/cuda/neural/neural_bvh_renderer_kernels.cuh:81-110; divide by w to display).
 * Per depth: closest-hit query over the live paths (hybrid with `environment` when it is not
 * NULL: its triangle meshes are diffuse with environment_albedo), then one shading kernel. A path
 * that escapes adds throughput * background and ends; a hit at max_depth ends; otherwise the
 * throughput is multiplied by the hit's albedo (the network's albedo for neural hits), Russian
 * roulette may end the path, and the next ray leaves in a cosine-distributed direction about the
 * shading normal. Live paths are compacted on the device. Every path writes its pixel exactly
 * once, so the frame is deterministic. Random numbers: the raygen LCG with frame = depth and
 * linear = pixel (two draws for the direction, one for Russian roulette), independent of the
 * compaction order. Primary rays: nbvh_raygen with (sample, frame 0).
 * The live-path counts stay on the device: every depth's launches are sized for width*height
 * rays and process the live ones, so the frame does not synchronise `stream` (the reference reads
 * the count back every step, src/cuda/neural/path/path_module.cu:124-129). With a workspace the
 * call allocates nothing and may be captured into a HIP graph (stats NULL; replays rerun the
 * frame with the parameters captured). stats may be NULL; when given, the call reads the counts
 * back once at the end (one synchronisation). bounce_query_flags with NBVH_QUERY_WAVEFRONT read
 * each depth's count back instead (no capture). */
int32_t nbvh_path_trace(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* environment,
                        const nbvh_path_params* params, const nbvh_camera* camera, uint32_t width, uint32_t height,
                        float* d_accum, nbvh_path_stats* stats, void* stream);

/* Bytes of device memory a frame of nbvh_path_trace with these arguments needs as its workspace
 * (about 180 B per pixel, more for hybrid scenes and sorted queries). No counterpart in the
 * reference, which sizes its path buffers at renderer set-up
 * (src/cuda/neural/path/path_module.cu:139-245). */
int32_t nbvh_path_trace_workspace_size(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* environment,
                                       const nbvh_path_params* params, uint32_t width, uint32_t height,
                                       uint64_t* bytes);

/* ---------------------------------------------------------------------------------------------
 * Training: the replacement of the reference's trainer->training_step(stream, inputs[9 x n],
 * targets[8 x n]) (src/cuda/neural/neural_bvh_renderer.cu:170-173) for its default
 * configuration, in FP32:
 *   loss   the composite loss of include/cuda/neural/utils/pt_composite_loss.cuh:19-123 with the
 *          defaults of include/cuda/neural/neural_bvh_module.h:250-266: visibility binary cross
 *          entropy; depth L1, normal L1 on the linearised output, albedo relative L2 (denominator
 *          held constant, include/cuda/neural/utils/loss_gradient.cuh), these three masked where
 *          the target visibility is < 0.5; each element divided by n * 8;
 *   update Adam with L2 regularisation added to the gradient (defaults of
 *          include/cuda/neural/utils/tcnn_config.h:153-166), under the reference's exponential
 *          learning-rate decay (src/cuda/neural/neural_bvh_module.cu:146-150: decay_start 10000,
 *          decay_interval 2000, decay_base 0.33): after each update number s >= decay_start with
 *          (s - decay_start) % decay_interval == 0 the rate is multiplied by decay_base, so update
 *          t uses learning_rate * decay_base^k, k = the number of such s < t. decay_interval 0
 *          turns the decay off. Parameters whose gradient is exactly zero (table entries no
 *          sample touched) are left unchanged.
 * The segment sampling and tree-cut learning of the reference's training loop are not part of
 * this: inputs are encoded segments (as for nbvh_network_inference) and targets are the 8 outputs
 * (nbvh_segment_labels gives them from the scene's triangles). The trainer holds its own FP32 copy
 * of the scene's network; nbvh_trainer_commit writes it back into the scene (and re-derives the
 * FP16/BF16/FP8 forms). Shapes: the MFMA query's, 96 inputs, 4 x 64 hidden, 8 outputs.
 * ------------------------------------------------------------------------------------------- */
typedef struct nbvh_trainer nbvh_trainer;

typedef struct nbvh_train_config {
    float    learning_rate; /* 1e-2 */
    float    beta1;         /* 0.9 */
    float    beta2;         /* 0.999 */
    float    epsilon;       /* 1e-8 */
    float    l2_reg;        /* 1e-8 */
    uint32_t max_batch;     /* samples per step, at most 2^24 (device scratch per sample: 720
                               floats, plus n_points * n_levels * 8 corner terms of 1 + F floats,
                               4.5 KB at 3 x 16 x 2; plus about 20 MB per trainer); and
                               max_batch * n_points * n_levels * 8 below 2^32 (11.1 M samples at
                               3 x 16 levels) */
    uint32_t decay_start;   /* 10000 in the reference; 0 with decay_interval 0: no decay */
    uint32_t decay_interval;/* 2000 in the reference; 0: no decay */
    float    decay_base;    /* 0.33 in the reference */
    uint32_t reserved;
} nbvh_train_config;

enum nbvh_train_flags {
    NBVH_TRAIN_GRADIENTS_ONLY = 1u /* compute loss and gradients, leave the parameters */
};

enum nbvh_train_array {
    NBVH_TRAIN_TABLES      = 0, /* the trained tables (tcnn order) */
    NBVH_TRAIN_MLP         = 1, /* the trained matrices (row-major, nbvh_network_init_params order) */
    NBVH_TRAIN_GRAD_TABLES = 2, /* the last step's gradients */
    NBVH_TRAIN_GRAD_MLP    = 3
};

int32_t nbvh_trainer_create(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_train_config* cfg, nbvh_trainer** out);
int32_t nbvh_trainer_destroy(nbvh_trainer* trainer);
/* One step on device buffers: d_inputs n x 9 (encoded segments), d_targets n x 8. Stream-ordered;
 * with loss != NULL the call waits for the step and returns its loss (sum of the element values). */
int32_t nbvh_trainer_step(nbvh_trainer* trainer, const float* d_inputs, const float* d_targets, uint32_t n,
                          uint32_t flags, float* loss, void* stream);
/* Host copy of one of the trainer's arrays (nbvh_train_array); out NULL: only *count. Synchronous. */
int32_t nbvh_trainer_get(nbvh_trainer* trainer, uint32_t what, float* out, uint64_t capacity, uint64_t* count);
/* Write the trained network into the scene. Synchronous; no query may run on the scene meanwhile. */
int32_t nbvh_trainer_commit(nbvh_trainer* trainer);

/* LDR-FLIP error of a rendering against a reference rendering, the image metric of the
 * reference's validation mode (compute_and_display_flip_error,
 * src/cuda/neural/neural_bvh_renderer.cu:40-92: both accumulation buffers divided by their sample
 * counts, linear RGB converted to sRGB, FLIP, the mean of the error map). d_reference / d_test:
 * width*height float4 accumulation buffers (rgb sum, sample count; as nbvh_path_trace writes them).
 * The metric follows the published method (Andersson et al., "FLIP: A Difference Evaluator for
 * Alternating Images", HPG 2020); sRGB values are clamped to [0, 1] (LDR-FLIP's input range) and
 * image borders replicate the edge pixels. ppd: pixels per degree, 0 => the reference's 0.7 m wide
 * 3840-pixel monitor at 0.7 m (67.02); at most about 118. d_error_map (optional): width*height
 * floats. *mean_error: the mean error. Synchronises `stream`; transient stream-ordered memory of
 * 32 B per pixel. */
int32_t nbvh_flip(nbvh_ctx* ctx, const float* d_reference, const float* d_test, uint32_t width, uint32_t height,
                  float ppd, float* d_error_map, float* mean_error, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Tree-cut learning: the reference's training loop (src/cuda/neural/neural_bvh_renderer.cu:
 * 141-297) around the trainer. Each step samples a batch of training segments from the N-BVH
 * being learned (granular_sample_bvh_network_input_data, src/cuda/neural/
 * neural_bvh_module_kernels.cu:119-264), runs one training step on them, adds every sample's loss
 * to its sampled node (update_neural_tree_cut_loss, src/cuda/neural/neural_bvh_renderer_kernels.cu:
 * 202-230), picks the leaves of largest loss heuristic (update_neural_tree_cut_decisions, :232-288)
 * and splits as many of them as the split schedule asks (BvhSplitScheduler, include/cuda/neural/
 * utils/bvh_split_scheduler.h:13-144; split_neural_bvh_at, src/cuda/neural/
 * neural_bvh_renderer_accel.cpp:256-392), each split adding the two children of the node in the
 * scene's triangle BVH; LoD assignment follows the schedule when more than one LoD is learned
 * (assign_lod_to_current_neural_leaf_nodes, :395-440). The learner owns the N-BVH: creating it
 * replaces the scene's N-BVH by the cut of its triangle BVH at neural_bvh_min_depth
 * (build_and_upload_neural_bvh, :505-558), and every split is written back to the scene.
 *
 * Sampler semantics kept from the reference: one thread per 2^log2_bvh_max_traversals samples;
 * its LCG is seeded per sample by get_lcg_rng(step * batch_size, 0, sample); a ray from a uniform
 * point of the scene box scaled by sampling_bbox_scaling in a uniform direction is traversed to
 * its next leaf at a LoD drawn among those learned, with boxes dilated by fixed_node_dilation +
 * U * node_extent_dilation and the learning epsilons; a segment without a triangle continues that
 * ray with the thread's next sample; with training_node_survival a sample is kept with
 * probability max(min_rr_survival_prob, heuristic / previous max heuristic) (1 above it), and a
 * dropped sample leaves its slot's inputs and targets as they were (the slot's node index is still
 * the newly sampled one, as in the reference). Not supported (NBVH_ERR_UNSUPPORTED): training
 * bounces (max_training_bounces > 1), view-dependent and leaf-boundary sampling, all off in the
 * reference's shipped configurations. Node losses are summed per step in FP64 before they are
 * added to the FP32 node loss (the reference adds FP32 atomics in an unspecified order); among
 * leaves of equal heuristic the lowest node index is split (the reference keeps one of them, the
 * last writer of a race).
 * ------------------------------------------------------------------------------------------- */
typedef struct nbvh_learner nbvh_learner;

typedef struct nbvh_learner_config {
    uint32_t batch_size;               /* samples per step (m_batch_size), 1..2^24; 2^18 in the
                                          shipped configurations */
    uint32_t log2_bvh_max_traversals;  /* samples per sampler thread = 2^this (0) */
    float    sampling_bbox_scaling;    /* 1.5 */
    float    fixed_node_dilation;      /* 1e-5 */
    float    node_extent_dilation;     /* 1e-4 */
    float    learning_mint_epsilon;    /* 1e-3 */
    float    learning_maxt_epsilon;    /* 1e-3 */
    uint32_t max_training_bounces;     /* 0 or 1 (1) */
    uint32_t training_node_survival;   /* 1 */
    float    survival_probability_scale; /* 1 */
    float    min_rr_survival_prob;     /* 0.005 */
    uint32_t tree_cut_learning;        /* 1: split leaves on the schedule */
    uint32_t neural_bvh_min_depth;     /* depth of the initial cut (0: the root alone) */
    uint32_t num_lods_learned;         /* 1..7 (1) */
    uint32_t max_training_steps;       /* steps planned (the schedule's split statistics) */
    uint32_t lod_decision;             /* 0 EqualIterations, 1 EqualSplits */
    int32_t  splitting_interval;       /* 50 */
    int32_t  last_split_iteration;     /* 2000 */
    int32_t  start_split_num;          /* 1 */
    float    split_scaling;            /* 2 */
    float    split_interval_scaling;   /* 2 */
    uint32_t reserved[4];
} nbvh_learner_config;

typedef struct nbvh_learner_state {
    uint32_t step;                /* steps done */
    uint32_t n_nodes;             /* N-BVH nodes */
    uint32_t current_learned_lod;
    uint32_t all_lods_assigned;
    uint32_t tree_cut_learning;   /* still splitting */
    uint32_t current_num_splits;  /* split candidates picked per step */
    uint32_t total_splits;        /* nodes split so far */
    uint32_t sample_failures;     /* sampler threads that found no leaf within 65536 rays */
    float    max_node_loss;       /* largest leaf heuristic after the last step */
    float    last_loss;           /* the last step's training loss */
    uint32_t expected_total_splits; /* the schedule's total over max_training_steps */
    uint32_t reserved[5];
} nbvh_learner_state;

/* Tree-cut data of one node (NeuralTreeCutData, include/cuda/neural/neural_bvh_renderer.h:36-47). */
typedef struct nbvh_tree_cut_node {
    float    loss;
    uint32_t iter_alive;
    uint32_t depth;
    uint32_t samples;
} nbvh_tree_cut_node;

enum nbvh_learner_array {
    NBVH_LEARN_INPUTS    = 0, /* float, batch x 9: the last sampled batch's encoded segments */
    NBVH_LEARN_TARGETS   = 1, /* float, batch x 8 */
    NBVH_LEARN_NODES     = 2, /* uint32, batch: each slot's sampled node */
    NBVH_LEARN_TREE_CUT  = 3, /* nbvh_tree_cut_node, one per N-BVH node */
    NBVH_LEARN_SPLITS    = 4, /* uint32, current_num_splits: the last step's split candidates
                                 (0 where fewer leaves had a positive heuristic) */
    NBVH_LEARN_NBVH      = 5  /* nbvh_node, the N-BVH as learned (boxes not dilated) */
};

/* Fills cfg with the reference's defaults (include/cuda/neural/neural_bvh_renderer_constants.cuh,
 * bvh_split_scheduler.h) and batch_size 2^18, max_training_steps 10000. */
int32_t nbvh_learner_config_defaults(nbvh_learner_config* cfg);
/* The scene needs its triangles (the targets and the splits come from its triangle BVH). */
int32_t nbvh_learner_create(nbvh_ctx* ctx, nbvh_scene* scene, const nbvh_learner_config* cfg, nbvh_learner** out);
int32_t nbvh_learner_destroy(nbvh_learner* learner);
/* Sample the batch of training step `step` with the learner's current N-BVH, tree-cut data and
 * max heuristic into its buffers (nbvh_learner_get); stream-ordered. */
int32_t nbvh_learner_sample(nbvh_learner* learner, uint32_t step, void* stream);
/* The tree-cut part of a step on the last sampled batch, given each slot's loss (d_sample_loss,
 * batch floats, device): node losses, leaf ageing and heuristics, split candidates, splits and LoD
 * assignment on the schedule; then the learner's step counter advances. Synchronises `stream`. */
int32_t nbvh_learner_update(nbvh_learner* learner, const float* d_sample_loss, void* stream);
/* One whole step: sample, train (trainer of the same scene), update. loss may be NULL. */
int32_t nbvh_learner_step(nbvh_learner* learner, nbvh_trainer* trainer, float* loss, void* stream);
int32_t nbvh_learner_get_state(const nbvh_learner* learner, nbvh_learner_state* state);
/* Host copy of a learner array (nbvh_learner_array) into out (capacity in elements of its type);
 * out NULL: only *count. Synchronous. */
int32_t nbvh_learner_get(nbvh_learner* learner, uint32_t what, void* out, uint64_t capacity, uint64_t* count);

/* Same with host buffers; synchronous. */
int32_t nbvh_query_host(nbvh_ctx* ctx, const nbvh_scene* scene, const nbvh_query_params* params,
                        const nbvh_ray* rays, uint32_t n_rays, nbvh_hit* hits);

/* Network inference: inputs [n x 3P] encoded segment points (in[i*3P + 3k + c]),
 * outputs [n x n_outputs] post-sigmoid (out[i*n_outputs + j]); optional raw pre-sigmoid
 * outputs (d_logits may be NULL). Replaces tcnn::Network::inference. */
int32_t nbvh_network_inference(nbvh_ctx* ctx, const nbvh_scene* scene, uint32_t precision, const float* d_inputs,
                               uint32_t n, float* d_outputs, float* d_logits, void* stream);

/* Host-pointer variant of nbvh_network_inference; synchronous. */
int32_t nbvh_network_inference_host(nbvh_ctx* ctx, const nbvh_scene* scene, uint32_t precision, const float* inputs,
                                    uint32_t n, float* outputs, float* logits);

/* ---------------------------------------------------------------------------------------------
 * Serving: continuous batching of ray-query requests over device-resident scenes.
 *
 * Replaces the reference's one-frame-at-a-time driver (NeuralPathModule::neural_bvh_traversal,
 * src/cuda/neural/path/path_module.cu:54-137, a single host thread with blocking syncs) for
 * callers that issue many independent ray batches. Any thread may submit; a scheduler thread
 * keeps up to n_streams launches in flight. Each launch takes the oldest queued request and
 * every later queued request with the same scene, kind family (neural or triangle; closest and
 * any hit mix) and query parameters (precision, lod, max_traversal_steps, visibility_threshold),
 * up to 64 requests and max_batch_rays rays, and runs them as one persistent kernel launch. A request's hits do not depend on what it was
 * batched with. Requests with host buffers are staged through a device arena of staging_bytes
 * (80 B per ray) and are admitted when their share fits; device-buffer requests are zero-copy.
 * The scene must not be modified while requests on it are queued or running.
 * ------------------------------------------------------------------------------------------- */
typedef struct nbvh_server nbvh_server;

typedef struct nbvh_server_config {
    uint32_t n_streams;      /* launches in flight at once, one HIP stream each, 1..16; 0 => 4 */
    uint32_t max_batch_rays; /* rays one launch collects from queued requests (a single larger
                                request still runs, alone); 0 => 2^21 */
    uint64_t staging_bytes;  /* device arena for host-buffer requests; 0 => 256 MiB */
    uint32_t reserved[4];
} nbvh_server_config;

enum nbvh_request_flags {
    NBVH_REQUEST_DEVICE_BUFFERS = 1u /* rays / hits are device pointers (no staging, no copies) */
};

typedef struct nbvh_request_info {
    int32_t  status;      /* NBVH_OK when the hits are complete */
    uint32_t n_launches;  /* launches the request ran in (1) */
    uint32_t batch_parts; /* requests in its launch, itself included */
    uint32_t reserved;
    uint64_t submit_ns;   /* steady-clock timestamps (ns): submitted, */
    uint64_t launch_ns;   /* its launch queued on a stream, */
    uint64_t complete_ns; /* hits complete (copied back for host buffers) */
} nbvh_request_info;

typedef struct nbvh_server_stats {
    uint64_t n_submitted;
    uint64_t n_completed;
    uint64_t n_failed;
    uint64_t n_launches;
    uint64_t n_batched_requests; /* sum over launches of requests per launch */
    uint64_t n_rays;             /* rays submitted */
    uint32_t n_pending;          /* queued, not launched */
    uint32_t n_inflight;         /* launches on the device */
} nbvh_server_stats;

int32_t nbvh_server_create(nbvh_ctx* ctx, const nbvh_server_config* cfg, nbvh_server** out_server);
/* Runs every queued request to completion, then frees the server. */
int32_t nbvh_server_destroy(nbvh_server* server);
/* Queue a request; *ticket identifies it for nbvh_request_wait. The buffers must stay valid until
 * the request completes. NBVH_ERR_CAPACITY: a host-buffer request larger than the arena. */
int32_t nbvh_request_submit(nbvh_server* server, const nbvh_scene* scene, const nbvh_query_params* params,
                            const nbvh_ray* rays, uint32_t n_rays, nbvh_hit* hits, uint32_t flags,
                            uint64_t* ticket);
/* Wait for a request (timeout_us < 0: no limit). NBVH_ERR_TIMEOUT leaves the ticket valid; any
 * other return collects it. info (optional) receives its timestamps and outcome. */
int32_t nbvh_request_wait(nbvh_server* server, uint64_t ticket, int64_t timeout_us, nbvh_request_info* info);
int32_t nbvh_server_get_stats(nbvh_server* server, nbvh_server_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* NBVH_H_ */
