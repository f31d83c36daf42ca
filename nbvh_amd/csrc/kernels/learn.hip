// learn.hip -- training data and tree-cut learning of the N-BVH (the reference's loss-driven
// sampling and splitting loop, src/cuda/neural/neural_bvh_renderer.cu:141-297).
//
//   learn_sample_kernel       granular_sample_bvh_network_input_data
//                             (src/cuda/neural/neural_bvh_module_kernels.cu:119-264) with
//                             stack_based_sample_uniform_bvh_density
//                             (include/cuda/neural/neural_bvh_module_common.cuh:703-780): one
//                             thread per sub-batch, rays from uniform points of the scaled scene
//                             box in uniform directions, traversed to their next accepted leaf with
//                             a per-sample dilation and the learning epsilons; optional
//                             loss-driven survival of the sampled node; encoded segment points and
//                             the triangle targets of the segment.
//   treecut_accumulate_kernel update_neural_tree_cut_loss
//                             (src/cuda/neural/neural_bvh_renderer_kernels.cu:202-230): each
//                             sample's loss values go to its sampled node.
//   treecut_update_kernel     the first pass of update_neural_tree_cut_decisions (:232-288): the
//                             node sums are added, leaves age by one iteration and every
//                             splittable leaf's loss heuristic becomes its sort key.
// The split candidates are then the leaves of largest heuristic (a descending radix sort, stable,
// so equal heuristics keep ascending node order); the host picks them and splits.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "../device/sampling.h"
#include "../device/triangles.h"
#include "../device/wave_query.h"
#include "launch.h"

namespace nbvh {
namespace dev {

// N-BVH nodes dilated on load by (1 + dil) about their centre (NeuralBVH2Node::
// intersect_primitive_min_max / intersect_traversal_bbox with a sampled dilation,
// include/cuda/neural/accel/neural_bvh2.cuh:69-117): centre = min + extent * 0.5.
struct DilatedNodes {
    const nbvh_node* nodes;
    float dil;
    __device__ __forceinline__ Box operator()(uint32_t i) const
    {
        Box b = load_node(nodes, i);
        const float f = 1.f + dil;
        const float ex = b.mxx - b.mnx, ey = b.mxy - b.mny, ez = b.mxz - b.mnz;
        const float cx = b.mnx + ex * 0.5f, cy = b.mny + ey * 0.5f, cz = b.mnz + ez * 0.5f;
        const float hx = ex * f, hy = ey * f, hz = ez * f;
        b.mnx = cx - hx * 0.5f;
        b.mxx = cx + hx * 0.5f;
        b.mny = cy - hy * 0.5f;
        b.mxy = cy + hy * 0.5f;
        b.mnz = cz - hz * 0.5f;
        b.mxz = cz + hz * 0.5f;
        return b;
    }
};

// NeuralTreeCutData::compute_loss_heuristic (include/cuda/neural/neural_bvh_renderer.h:42-46).
__device__ __forceinline__ float loss_heuristic(const TreeCutNode& t)
{
    return t.samples == 0u ? -FLT_MAX : t.loss / ((float)t.iter_alive * sqrtf((float)t.samples / (float)t.iter_alive));
}

// A sampler thread gives up on a sample after this many rays that found no leaf (the reference
// retries without bound; with the scene box sampled this never happens in practice).
constexpr uint32_t kMaxSampleAttempts = 1u << 16;

__global__ __launch_bounds__(256) void learn_sample_kernel(SampleArgs a)
{
    __shared__ uint32_t stack_lds[4 * kNeuralStackSize * 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    LdsStack st{stack_lds + wave * kNeuralStackSize * 64 + lane, 0};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.sub_batch_size) return;  // no barrier follows
    const uint32_t n_in = 3u * a.n_points;
    Ray ray{0.f, 0.f, 0.f, 1e-4f, 1.f, 0.f, 0.f, FLT_MAX};
    bool continue_traversal = false;
    uint32_t thread_idx = i * a.n_sub_batch;
    const uint32_t n_lods = NBVH_NUM_LODS - a.current_lod;
    for (uint32_t inner = 0; inner < a.n_sub_batch; inner++) {
        uint32_t rng = lcg_seed(a.step * a.batch_size, 0u, thread_idx);
        // stack_based_sample_uniform_bvh_density: a direction, a LoD among those learned, a dilation.
        float dir[3];
        {
            const float u0 = lcg_float(rng), u1 = lcg_float(rng);
            const float phi = 6.28318548202514648f * u0;
            const float ct = u1 * 2.f - 1.f;
            const float stt = sqrtf(1.f - ct * ct);
            float sp, cph;
            sincosf_cw(phi, sp, cph);
            dir[0] = stt * cph;
            dir[1] = stt * sp;
            dir[2] = ct;
        }
        const uint32_t k = (uint32_t)(lcg_float(rng) * (float)n_lods);
        const uint32_t lod = a.current_lod + (k < n_lods - 1u ? k : n_lods - 1u);
        const float dil = a.fixed_dilation + lcg_float(rng) * a.extent_dilation;
        bool valid = continue_traversal && !st.empty();
        float mint = 0.f, maxt = 0.f;
        uint32_t node = 0;
        uint32_t attempts = 0;
        const DilatedNodes nodes{a.nodes, dil};
        while (true) {
            if (!valid) {
                st.n = 0;
                st.push(0u);
                const float h = (a.scaling - 1.f) * 0.5f;
                const float px = a.scaling * lcg_float(rng) - h;
                const float py = a.scaling * lcg_float(rng) - h;
                const float pz = a.scaling * lcg_float(rng) - h;
                ray = Ray{px * a.box_ext[0] + a.box_min[0], py * a.box_ext[1] + a.box_min[1],
                          pz * a.box_ext[2] + a.box_min[2], 1e-4f, dir[0], dir[1], dir[2], FLT_MAX};
            }
            const int r = nbvh_next_leaf(nodes, lod, ray, st, FLT_MAX, mint, maxt, node, 0xFFFFFFFFu, a.mint_eps,
                                         a.maxt_eps);
            valid = r == 1;
            if (valid || ++attempts >= kMaxSampleAttempts) break;
        }
        if (!valid) {
            atomicAdd(a.failures, 1u);
            st.n = 0;
            continue_traversal = false;
            continue;
        }
        a.node_idx[thread_idx] = node;
        if (a.survival) {  // loss-driven survival of the sampled node
            const float est = loss_heuristic(a.tree_cut[node]);
            float keep = fmaxf(a.min_survival, est / a.current_max_loss);
            if (!(est <= a.current_max_loss)) keep = 1.f;
            if (a.survival_scale * keep < lcg_float(rng)) continue;
        }
        // encode_segment_from_world_po_pi (neural_bvh_module_kernels.cu:10-97): points from the
        // segment's entry to its exit, no swap for a negative entry.
        {
            const float po[3] = {ray.ox + mint * ray.dx, ray.oy + mint * ray.dy, ray.oz + mint * ray.dz};
            const float pi[3] = {ray.ox + maxt * ray.dx, ray.oy + maxt * ray.dy, ray.oz + maxt * ray.dz};
            float eo[3], ei[3];
            for (int c = 0; c < 3; c++) {
                eo[c] = (po[c] - a.enc_center[c]) / a.enc_extent[c] + 0.5f;
                ei[c] = (pi[c] - a.enc_center[c]) / a.enc_extent[c] + 0.5f;
            }
            float* in = a.inputs + (size_t)thread_idx * n_in;
            for (uint32_t p = 0; p < a.n_points; p++) {
                const float t = a.seg_t[p], w = 1.f - t;
                for (int c = 0; c < 3; c++) in[3 * p + c] = w * eo[c] + t * ei[c];
            }
        }
        float out[8];
        const bool hit = segment_targets(a.tri, ray, mint, maxt, out);
        for (int j = 0; j < 8; j++) a.targets[(size_t)thread_idx * 8 + j] = out[j];
        // A segment without a hit continues the traversal of its ray with the next sample; with at
        // most one training bounce a hit ends it (max_training_bounces 0 or 1).
        continue_traversal = !hit;
        thread_idx++;
    }
}

__global__ __launch_bounds__(256) void treecut_accumulate_kernel(const uint32_t* __restrict__ node_idx,
                                                                 const float* __restrict__ sample_loss, uint32_t n,
                                                                 double* __restrict__ sums,
                                                                 uint32_t* __restrict__ counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t node = node_idx[i];
    atomicAdd(sums + node, (double)sample_loss[i]);
    atomicAdd(counts + node, 8u);  // one sample per output element (8 outputs)
}

__global__ __launch_bounds__(256) void treecut_update_kernel(const nbvh_node* __restrict__ nodes, uint32_t n_nodes,
                                                             TreeCutNode* __restrict__ tc, double* __restrict__ sums,
                                                             uint32_t* __restrict__ counts,
                                                             float* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    TreeCutNode t = tc[i];
    const uint32_t c = counts[i];
    if (c != 0u) {
        t.loss = t.loss + (float)sums[i];
        t.samples += c;
        sums[i]   = 0.0;
        counts[i] = 0u;
    }
    float key = 0.f;  // not a split candidate
    const Box b = load_node(nodes, i);
    if (b.left == NBVH_NO_NEURAL_CHILD) {
        t.iter_alive += 1u;
        if (b.prim_count != NBVH_LEAF_FINAL) {
            const float h = loss_heuristic(t);
            key = h > 0.f ? h : 0.f;
        }
    }
    tc[i]   = t;
    keys[i] = key;
    vals[i] = i;
}

}  // namespace dev

hipError_t learn_sample(const SampleArgs& a, hipStream_t stream)
{
    const uint32_t blocks = (a.sub_batch_size + 255u) / 256u;
    if (blocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(dev::learn_sample_kernel, dim3(blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
}

size_t treecut_sort_scratch_bytes(uint32_t n_nodes)
{
    size_t tmp = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, (const float*)nullptr, (float*)nullptr,
                                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n_nodes);
    return tmp;
}

hipError_t treecut_update(const uint32_t* node_idx, const float* sample_loss, uint32_t n, const nbvh_node* nodes,
                          uint32_t n_nodes, TreeCutNode* tc, double* sums, uint32_t* counts, float* keys,
                          uint32_t* vals, float* keys_sorted, uint32_t* vals_sorted, void* tmp, size_t tmp_bytes,
                          hipStream_t stream)
{
    if (n > 0u) {
        hipLaunchKernelGGL(dev::treecut_accumulate_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, node_idx,
                           sample_loss, n, sums, counts);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dev::treecut_update_kernel, dim3((n_nodes + 255u) / 256u), dim3(256), 0, stream, nodes,
                       n_nodes, tc, sums, counts, keys, vals);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, keys, keys_sorted, vals, vals_sorted,
                                                        (int)n_nodes, 0, 32, stream);
}

}  // namespace nbvh
