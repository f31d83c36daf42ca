// launch.h -- host-callable launchers of the HIP kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>

#include "../nbvh_internal.h"

namespace nbvh {

// Primary rays (camera_raygen with LCG jitter).
hipError_t raygen(const nbvh_camera& cam, uint32_t w, uint32_t h, uint32_t sample, uint32_t frame, nbvh_ray* rays,
                  hipStream_t stream);

// Zero a small device counter array (a kernel, not hipMemsetAsync, so the call can be captured).
hipError_t zero_u32(uint32_t* p, uint32_t n, hipStream_t stream);

// One diffuse bounce: next rays from rays + hits (tri.verts may be null when no triangle hits occur).
hipError_t scatter_diffuse(const TriArgs& tri, const nbvh_ray* rays, const nbvh_hit* hits, uint32_t n, uint32_t sample,
                           uint32_t bounce, nbvh_ray* out, hipStream_t stream);

// Network targets (labels) of segments from the triangle BVH: targets[n x 8].
hipError_t segment_labels(const TriArgs& a, const nbvh_ray* rays, const float* mint, const float* maxt, uint32_t n,
                          float* targets, hipStream_t stream);

// Coherence order of a ray batch (sorted by part, direction octant, origin Morton code over the
// box enc_min + [0, enc_ext]): order[claim] = ray index. Scratch: ray_order_scratch_bytes(n).
size_t ray_order_scratch_bytes(uint32_t n);
hipError_t ray_order(const RayBatch& rb, const float enc_min[3], const float enc_ext[3], uint32_t* order,
                     void* scratch, size_t scratch_bytes, hipStream_t stream);
// The same grouping from kernels alone (recordable into a HIP graph): 2^16 buckets of (octant,
// coarse origin Morton code), order[claim] for the claims [0, count of rb); no order inside a
// bucket. Scratch: ray_bucket_scratch_bytes(n).
size_t ray_bucket_scratch_bytes(uint32_t n);
hipError_t ray_bucket_order(const RayBatch& rb, const float enc_min[3], const float enc_ext[3], uint32_t* order,
                            void* scratch, hipStream_t stream);

// Hybrid query steps: rays clipped to their triangle hits; neural records merged over triangle ones.
// live: optional device-side count (at most n) of the rays to process.
hipError_t hybrid_clip(const nbvh_ray* rays, const nbvh_hit* tri, uint32_t n, const uint32_t* live, nbvh_ray* out,
                       hipStream_t stream);
hipError_t hybrid_merge(nbvh_hit* hits, const nbvh_hit* neural, uint32_t n, const uint32_t* live, bool any_hit,
                        hipStream_t stream);

// Sorted wavefront rounds (wavefront.hip): state init, traversal to the next leaf per active ray
// (segments keyed by the Morton code of their midpoints), segment sort by key, encoding in key
// order, and the consume step after the network.
hipError_t wf_init(const RayBatch& b, WfRayState* st, uint32_t* stacks, uint32_t* active, hipStream_t stream);
hipError_t wf_traverse(const NetArgs& a, const RayBatch& b, WfRayState* st, uint32_t* stacks, const uint32_t* active,
                       uint32_t n_active, WfSeg* segs, uint32_t* keys, uint32_t* vals, uint32_t* counter,
                       hipStream_t stream);
size_t wf_sort_scratch_bytes(uint32_t n);
hipError_t wf_sort(void* tmp, size_t tmp_bytes, const uint32_t* keys, uint32_t* keys_out, const uint32_t* vals,
                   uint32_t* order, uint32_t m, int end_bit, hipStream_t stream);
hipError_t wf_encode(const NetArgs& a, const RayBatch& b, const WfSeg* segs, const uint32_t* order, uint32_t m,
                     float* in, hipStream_t stream);
hipError_t wf_consume(const NetArgs& a, const RayBatch& b, const WfSeg* segs, const uint32_t* order,
                      const float* logits, uint32_t m, WfRayState* st, uint32_t* next, uint32_t* next_count,
                      hipStream_t stream);

// Training (train.hip): the activation row stride and workspace bytes of a max batch, zero fill, loss + gradients of one batch of
// encoded inputs (9 floats) and targets (8 floats) for the supported shapes (*supported), Adam.
uint32_t train_act_stride(uint32_t max_batch);
size_t   train_act_bytes(uint32_t max_batch);
size_t   train_bin_bytes(uint32_t max_batch, uint32_t n_points, uint32_t n_levels, uint32_t n_features);
hipError_t train_zero(float* p, uint64_t n, hipStream_t stream);
hipError_t train_gradients(const TrainArgs& t, const float* in, const float* targets, uint32_t n, hipStream_t stream,
                           bool* supported);
hipError_t train_adam(float* w, const float* g, float* m, float* v, uint32_t* steps, uint64_t n, bool matrix,
                      const AdamArgs& h, hipStream_t stream);

// Path-trace driver: per-path state init (pixel = index, throughput 1) and one depth's shading
// and compaction (live paths appended at *counter) of min(*live, n) paths (n without live).
hipError_t path_init(PathState* st, uint32_t n, hipStream_t stream);
hipError_t path_shade(const PathArgs& a, uint32_t depth, uint32_t n, const uint32_t* live, const nbvh_ray* rays,
                      const nbvh_hit* hits, const PathState* st, nbvh_ray* next_rays, PathState* next_st,
                      uint32_t* counter, hipStream_t stream);

// Triangle BVH closest / any hit.
hipError_t query_triangles(const TriArgs& a, const RayBatch& rb, hipStream_t stream);

// Neural queries (persistent grid sized from n_cus and the kernel's occupancy).
// *supported = false when no kernel instance matches the network shape.
hipError_t query_fp32(const NetArgs& a, const RayBatch& rb, uint32_t* queue, uint32_t n_cus, hipStream_t stream,
                      bool* supported);
hipError_t network_fp32(const NetArgs& a, const float* in, uint32_t n, float* out, float* logits, hipStream_t stream,
                        bool* supported);

// MFMA paths: precision = NBVH_PREC_FP16 / BF16 / FP8.
hipError_t query_mfma(const NetArgs& a, uint32_t precision, const RayBatch& rb, uint32_t* queue, uint32_t n_cus,
                      hipStream_t stream, bool* supported);
// The network kernels are persistent too: at most n_cus x (resident workgroups per CU) blocks,
// each staging the weights once.
hipError_t network_mfma(const NetArgs& a, uint32_t precision, const float* in, uint32_t n, float* out, float* logits,
                        uint32_t n_cus, hipStream_t stream, bool* supported);

// Tree-cut learning (learn.hip): the training-data sampler; the per-node loss update, the leaves'
// loss heuristics and their descending sort (keys_sorted / vals_sorted: heuristic, node).
hipError_t learn_sample(const SampleArgs& a, hipStream_t stream);
size_t treecut_sort_scratch_bytes(uint32_t n_nodes);
hipError_t treecut_update(const uint32_t* node_idx, const float* sample_loss, uint32_t n, const nbvh_node* nodes,
                          uint32_t n_nodes, TreeCutNode* tc, double* sums, uint32_t* counts, float* keys,
                          uint32_t* vals, float* keys_sorted, uint32_t* vals_sorted, void* tmp, size_t tmp_bytes,
                          hipStream_t stream);

// LDR-FLIP (flip.hip): the error map (optional) and its mean (device double) of two accumulation
// buffers; flip_tiles(w, h) partial sums in a.partials.
uint32_t flip_tiles(uint32_t w, uint32_t h);
hipError_t flip_error(const FlipArgs& a, double* mean, hipStream_t stream);

// Static scratch of the allocation-free query (noalloc.hip): kNoAllocScratchBytes per device.
// noalloc_scratch_init looks up the array's device address on `device` (which must be current)
// once; noalloc_scratch returns it (NULL before the lookup).
constexpr size_t kNoAllocScratchBytes = 24u << 20;
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);

}  // namespace nbvh
