// wave_query.h -- the persistent wave scheduler shared by every neural query kernel.
//
// Reference structure replaced: NeuralPathModule::neural_bvh_traversal
// (src/cuda/neural/path/path_module.cu:54-137) runs one kernel per traversal step over all active
// rays, compacts the survivors with atomicAdd into a ping-pong buffer
// (path_inference_kernels.cu:152-193) and copies the active count to the host after every step.
// Here one wave owns 64 ray slots for the whole launch. Each iteration it
//   (A) refills empty slots from a global ray queue (one atomic per wave) and advances every
//       slot's N-BVH traversal (neural_bvh2.cuh:170-292) until the slot holds a leaf segment,
//       finishing rays whose stack is exhausted;
//   (B) evaluates the 64 segments together (encoding + hash grid + MLP, done by the policy);
//   (C) decodes the outputs per slot (neural_bvh_module_common.cuh:160-209,497-550,
//       path_inference_kernels.cu:61-131) and keeps the closest depth.
// The per-ray semantics are those of the reference's step loop: each step evaluates the next
// accepted leaf in traversal order with current_maxt = closest depth so far; at most
// max_steps + 1 segments per ray. No host round trip, no compaction traffic.
#pragma once

#include "common.h"

namespace nbvh {
namespace dev {

// Traversal stack in LDS, one column per lane (entry e of lane l at [e * 64 + l]).
struct LdsStack {
    uint32_t* col;
    int n;
    __device__ __forceinline__ void push(uint32_t v)
    {
        col[n * 64] = v;
        n++;
    }
    __device__ __forceinline__ uint32_t pop()
    {
        n--;
        return col[n * 64];
    }
    __device__ __forceinline__ bool empty() const { return n == 0; }
    __device__ __forceinline__ int size() const { return n; }
};

// EMPTY: needs a ray; NEW: ray claimed, not loaded; RAY: traversing to the next leaf;
// SEG: holds a leaf segment to evaluate; DONE: the queue is drained.
enum SlotState : uint32_t { SLOT_EMPTY = 0, SLOT_RAY = 1, SLOT_SEG = 2, SLOT_DONE = 3, SLOT_NEW = 4 };

// Per-ray result of a neural query, built up segment by segment (shared by the persistent wave
// scheduler and the sorted wavefront rounds, so both make identical decisions).
struct HitState {
    float best_t;  // closest decoded depth so far (starts at the ray's tmax)
    uint32_t segments, flags, kind, hit_node;
    float hit_logit;
    float nrm[3], alb[3];
};

__device__ __forceinline__ void hit_state_init(HitState& h, float tmax)
{
    h.best_t    = tmax;
    h.segments  = 0;
    h.flags     = 0;
    h.kind      = 0;
    h.hit_node  = 0xFFFFFFFFu;
    h.hit_logit = 0.f;
    h.nrm[0] = h.nrm[1] = h.nrm[2] = 0.f;
    h.alb[0] = h.alb[1] = h.alb[2] = 0.f;
}

// The hit record of a finished ray.
__device__ __forceinline__ nbvh_hit make_hit(const HitState& s, bool any_hit)
{
    nbvh_hit h;
    h.t       = (s.kind != 0u && !any_hit) ? s.best_t : FLT_MAX;
    h.prim_id = 0xFFFFFFFFu;
    h.node_id = s.kind ? s.hit_node : 0xFFFFFFFFu;
    h.info    = s.kind | (s.flags << 8) | ((s.segments < 65535u ? s.segments : 65535u) << 16);
    h.normal_or_uv[0] = s.nrm[0];
    h.normal_or_uv[1] = s.nrm[1];
    h.normal_or_uv[2] = s.nrm[2];
    h.logit     = s.kind ? s.hit_logit : 0.f;
    h.albedo[0] = s.alb[0];
    h.albedo[1] = s.alb[1];
    h.albedo[2] = s.alb[2];
    h.reserved  = 0.f;
    return h;
}

// One evaluated segment (outputs z[0..7], pre-sigmoid) applied to the ray's result
// (neural_bvh_module_common.cuh:160-209,497-550, path_inference_kernels.cu:61-131): visibility
// decision on the logit, LocalT depth decode, closest depth kept with its normal and albedo.
// Returns true when the ray is finished (an any-hit query found an occluding segment).
// Decoded hit distance of a segment's depth output y (post-sigmoid): the network's depth mode
// (LocalT on the segment, WorldT on the ray's span through the neural scene's box) and, with the
// linear activation, u = clamp((y - 0.45) * 10, 0, 1) (decode_depth_estimate,
// include/cuda/neural/neural_bvh_module_common.cuh:23-84). FLT_MAX = rejected.
__device__ __forceinline__ float decode_segment_depth(const NetArgs& a, const Ray& r, float y, float seg_mint,
                                                      float seg_maxt)
{
    const float u = a.depth_linear ? fminf(fmaxf((y - 0.45f) * 10.f, 0.f), 1.f) : y;
    if (a.depth_mode == NBVH_DEPTH_WORLD_T) {
        const Box b{a.bbox_min[0], a.bbox_min[1], a.bbox_min[2], a.bbox_max[0], a.bbox_max[1], a.bbox_max[2]};
        float smin, smax;
        slab(b, r, smin, smax);
        return decode_span(u, smin, smax, seg_mint < 0.f);
    }
    return decode_depth(u, seg_mint, seg_maxt);
}

__device__ __forceinline__ bool apply_segment(const NetArgs& a, HitState& s, bool any_hit, const Ray& r,
                                              float seg_mint, float seg_maxt, uint32_t seg_node, const float z[8])
{
    s.segments++;
    if (!(z[0] > a.thr_logit)) return false;
    const bool inside = seg_mint < 0.f;
    if (any_hit) {
        bool h = true;
        if (inside) h = decode_segment_depth(a, r, sigmoidf_cw(z[1]), seg_mint, seg_maxt) != FLT_MAX;
        if (h) {
            s.kind      = NBVH_HIT_NEURAL;
            s.hit_node  = seg_node;
            s.hit_logit = z[0];
        }
        return h;
    }
    const float depth = decode_segment_depth(a, r, sigmoidf_cw(z[1]), seg_mint, seg_maxt);
    if (depth == FLT_MAX) return false;
    // interpolate_depth: an estimate within interp_eps of the current neural hit is averaged with it
    // (process_neural_bvh_traversal_step, src/cuda/neural/path/path_inference_kernels.cu:89-121).
    const bool near = a.interpolate_depth && s.kind == NBVH_HIT_NEURAL && fabsf(s.best_t - depth) <= a.interp_eps;
    if (depth < s.best_t || near) {
        const float prev_t = s.best_t;
        const float prev_n[3] = {s.nrm[0], s.nrm[1], s.nrm[2]};
        s.best_t    = depth;
        s.kind      = NBVH_HIT_NEURAL;
        s.hit_node  = seg_node;
        s.hit_logit = z[0];
        float v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c]     = sigmoidf_cw(z[2 + c]) * 2.f - 1.f;
            s.alb[c] = sigmoidf_cw(z[5 + c]);
        }
        const float inv = 1.f / sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
#pragma unroll
        for (int c = 0; c < 3; c++) s.nrm[c] = v[c] * inv;
        if (near) {
            s.best_t = 0.5f * (prev_t + depth);
            float m[3];
#pragma unroll
            for (int c = 0; c < 3; c++) m[c] = s.nrm[c] + prev_n[c];
            const float im = 1.f / sqrtf((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]);
#pragma unroll
            for (int c = 0; c < 3; c++) s.nrm[c] = m[c] * im;
        }
    }
    return false;
}

struct Slot {
    uint32_t state;
    uint32_t part;    // part of a multi-part launch (0 for a single array)
    uint32_t ray_id;  // index within the part
    bool any_hit;     // stop at the first accepted segment (shadow ray)
    Ray ray;
    float seg_mint, seg_maxt;
    uint32_t seg_node;
    uint32_t rng;  // the ray's LCG stream (uniform segment distribution)
    HitState h;
};

template <bool PARTS>
__device__ __forceinline__ void finish_ray(const NetArgs& a, Slot& s, const RayBatch& b)
{
    store_hit(batch_hits<PARTS>(b, s.part), s.ray_id, make_hit(s.h, s.any_hit));
    s.state = SLOT_EMPTY;
}

// XCD (accelerator complex die) the wave runs on: HW_REG_XCC_ID bits [3:0].
__device__ __forceinline__ uint32_t xcc_id()
{
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu;
}

// Ray queue split into `bands` (at most kQueueBands) contiguous bands of the ray array, one
// counter per band.
// A wave starts on the band of its XCD, so rays that are neighbours in the array (neighbouring
// pixels for image-ordered rays) are traced on the same XCD and share its L2; a wave moves on
// to the next band only when its band is drained (work stealing for the tail).
#ifndef NBVH_QUEUE_BANDS
#define NBVH_QUEUE_BANDS 8
#endif
constexpr uint32_t kQueueBands   = NBVH_QUEUE_BANDS;
constexpr uint32_t kQueueStride  = 4;                             // counters 16 B apart
constexpr uint32_t kQueueWords   = kQueueBands * kQueueStride;    // u32 words per query

struct RayQueue {
    uint32_t* heads;
    uint32_t n_rays;
    uint32_t image_width;  // 0: claim order = array order; else 8x8-tile order of a scanline image
    uint32_t band;         // wave-uniform
    uint32_t tried;        // wave-uniform: bands found drained
    uint32_t bands;        // 1..kQueueBands (launch-uniform)
    __device__ __forceinline__ void init(uint32_t* h, uint32_t n, uint32_t width, uint32_t n_bands)
    {
        heads = h;
        n_rays = n;
        image_width = width;
        bands = (n_bands >= 1u && n_bands <= kQueueBands) ? n_bands : kQueueBands;
        band = xcc_id() % bands;
        tried = 0;
    }
    __device__ __forceinline__ uint32_t band_lo(uint32_t b) const
    {
        return (uint32_t)(((uint64_t)b * n_rays) / bands);
    }
    // Claim index -> ray index. With an image width W (W % 8 == 0, n % (8W) == 0, checked on the
    // host) consecutive claims walk 8x8 pixel tiles, tile rows of 8 scanlines in order, so the
    // 64 slots of a wave trace a compact pixel block (gathers of neighbouring rays share cache
    // lines at the coarse and middle grid levels). A bijection on [0, n).
    __device__ __forceinline__ uint32_t ray_of(uint32_t i) const
    {
        if (image_width == 0u) return i;
        const uint32_t block = i / (8u * image_width), j = i - block * 8u * image_width;
        const uint32_t tile = j >> 6, w = j & 63u;
        return (block * 8u + (w >> 3)) * image_width + tile * 8u + (w & 7u);
    }
};

// Phase A. Returns false when the wave has no segment left to evaluate (all slots DONE).
// Traversal rounds: every slot still looking for a leaf advances at most kTraversalBudget nodes
// per round; the wave goes on to evaluate its segments once at least kReadyNum / kReadyDen of its
// live slots hold one, so one long traversal does not hold up the other 63 slots' network
// evaluation (the suspended slots resume in the next iteration; results are unchanged).
#ifndef NBVH_TRAVERSAL_BUDGET
#define NBVH_TRAVERSAL_BUDGET 16
#endif
#ifndef NBVH_READY_NUM
#define NBVH_READY_NUM 1
#endif
#ifndef NBVH_READY_DEN
#define NBVH_READY_DEN 2
#endif
constexpr uint32_t kTraversalBudget = NBVH_TRAVERSAL_BUDGET;

template <bool PARTS, typename Nodes>
__device__ __forceinline__ bool fill_slots(const NetArgs& a, const Nodes& nodes, Slot& s, LdsStack& st,
                                           const RayBatch& b, RayQueue& q)
{
    const uint32_t lane = threadIdx.x & 63u;
    while (true) {
        // Claim rays for the empty slots (convergent: every lane takes part in the ballot).
        const uint64_t m = __ballot(s.state == SLOT_EMPTY);
        if (m != 0ull) {
            if (q.tried >= q.bands) {
                if (s.state == SLOT_EMPTY) s.state = SLOT_DONE;
            } else {
                const uint32_t leader = __ffsll((unsigned long long)m) - 1u;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(q.heads + q.band * kQueueStride, (uint32_t)__popcll(m));
                base = __shfl(base, leader);
                const uint32_t lo = q.band_lo(q.band), hi = q.band_lo(q.band + 1u);
                bool failed = false;
                if (s.state == SLOT_EMPTY) {
                    const uint32_t r = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                    if (r < hi - lo) {
                        s.state = SLOT_NEW;
                        batch_locate<PARTS>(b, batch_ray(b, q.ray_of(lo + r)), s.part, s.ray_id);
                    } else {
                        failed = true;
                    }
                }
                if (__any(failed)) {
                    q.band = (q.band + 1u) % q.bands;
                    q.tried++;
                }
            }
        }
        if (s.state == SLOT_NEW) {
            s.state    = SLOT_RAY;
            s.ray      = load_ray(batch_rays<PARTS>(b, s.part), s.ray_id);
            s.any_hit  = batch_any_hit<PARTS>(b, s.part, a.any_hit);
            s.rng      = a.seg_uniform ? lcg_seed(a.seg_seed, 0u, s.ray_id) : 0u;
            hit_state_init(s.h, s.ray.tmax);
            st.n = 0;
            st.push(0u);
        }
        if (s.state == SLOT_RAY) {
            if (st.empty()) {
                finish_ray<PARTS>(a, s, b);
            } else {
                const int r = nbvh_next_leaf(nodes, a.lod, s.ray, st, s.h.best_t, s.seg_mint, s.seg_maxt,
                                             s.seg_node, kTraversalBudget);
                if (r == 2) {
                    // suspended: stays SLOT_RAY
                } else if (r == 1) {
                    if (s.h.segments == a.max_steps + 1u) {
                        s.h.flags |= NBVH_FLAG_STEP_CAP;
                        finish_ray<PARTS>(a, s, b);
                    } else {
                        s.state = SLOT_SEG;
                    }
                } else {
                    if (r < 0) s.h.flags |= NBVH_FLAG_STACK_OVERFLOW;
                    finish_ray<PARTS>(a, s, b);
                }
            }
        }
        if (__all(s.state == SLOT_SEG || s.state == SLOT_DONE)) break;
        const uint32_t ready = (uint32_t)__popcll(__ballot(s.state == SLOT_SEG));
        const uint32_t live  = (uint32_t)__popcll(__ballot(s.state != SLOT_DONE));
        if (ready * NBVH_READY_DEN >= live * NBVH_READY_NUM && ready > 0u) break;
    }
    return !__all(s.state == SLOT_DONE);
}

// Phase C: outputs z[0..7] (pre-sigmoid logits) of the slot's current segment.
template <bool PARTS>
__device__ __forceinline__ void consume_segment(const NetArgs& a, Slot& s, const float z[8], const RayBatch& b)
{
    s.state = SLOT_RAY;
    if (apply_segment(a, s.h, s.any_hit, s.ray, s.seg_mint, s.seg_maxt, s.seg_node, z)) finish_ray<PARTS>(a, s, b);
}

}  // namespace dev
}  // namespace nbvh
