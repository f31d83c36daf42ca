// api.cpp -- the extern "C" boundary of libnbvh_hip.so (declared in include/nbvh.h).
//
// Mirrors the reference's NeuralBVHModule life cycle (include/cuda/neural/neural_bvh_module.h:144-241):
// scenes own device memory (the reference: tcnn GPUMemoryArena + CUDABuffer), N-BVH and weights
// are loaded per scene (load_neural_bvh / load_network_model), a query runs the traversal loop
// of NeuralPathModule::neural_bvh_traversal (src/cuda/neural/path/path_module.cu:54-137) as one
// persistent kernel. Errors are status codes with a message (the reference throws from
// CUDA_CHECK and calls exit(2) from CUDA_SYNC_CHECK, include/ntwr.h:30-54).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../kernels/launch.h"
#include "context.h"
#include "mfma_layout.h"

namespace nbvh {

thread_local std::string t_last_error;

void set_error(const nbvh_ctx* ctx, const std::string& msg)
{
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        const_cast<nbvh_ctx*>(ctx)->last_error = msg;
    }
    t_last_error = msg;
}

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(NBVH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void require(bool cond, const char* msg)
{
    if (!cond) fail(NBVH_ERR_INVALID_ARGUMENT, msg);
}

void DevBuf::release()
{
    if (ptr) {
        (void)hipSetDevice(ctx->device);
        (void)hipFree(ptr);
        ctx->used -= bytes;
    }
    ptr   = nullptr;
    bytes = 0;
}

void DevBuf::alloc(nbvh_ctx* c, size_t n)
{
    release();
    ctx = c;
    if (n == 0) return;
    if (ctx->used.load() + n > ctx->budget)
        fail(NBVH_ERR_OUT_OF_MEMORY, "HBM budget exceeded: " + std::to_string(ctx->used.load() + n) + " > " +
                                         std::to_string(ctx->budget) + " bytes");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (hipMalloc(&ptr, n) != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        fail(NBVH_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    ctx->used += n;
}

namespace {

// MLP layout consumed by the FP32 kernels: layers 1, 3, 5 row-major, layers 2, 4 transposed.
std::vector<float> fp32_kernel_layout(const nbvh_network_config& cfg, const std::vector<float>& mlp)
{
    const uint32_t K0 = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    std::vector<float> out(mlp.size());
    size_t off = 0;
    uint32_t in = K0;
    for (uint32_t layer = 0; layer <= cfg.n_hidden_layers; layer++) {
        const uint32_t o = layer == cfg.n_hidden_layers ? cfg.n_outputs : cfg.n_neurons;
        const bool transposed = (layer % 2 == 1) && layer != cfg.n_hidden_layers;
        for (uint32_t j = 0; j < o; j++)
            for (uint32_t k = 0; k < in; k++) {
                const float w = mlp[off + (size_t)j * in + k];
                if (transposed)
                    out[off + (size_t)k * o + j] = w;
                else
                    out[off + (size_t)j * in + k] = w;
            }
        off += (size_t)o * in;
        in = cfg.n_neurons;
    }
    return out;
}

}  // namespace

void refresh_args(nbvh_scene* s)
{
    NetArgs& a = s->args;
    a = NetArgs{};
    a.nodes   = s->d_nbvh.as<nbvh_node>();
    a.n_nodes = (uint32_t)s->nbvh_nodes.size();
    const float dil = (1.f + kNodeExtentDilation) + kFixedNodeDilation;
    for (int c = 0; c < 3; c++) {
        a.enc_center[c] = (s->bbox_max[c] + s->bbox_min[c]) * 0.5f;
        a.enc_extent[c] = (s->bbox_max[c] - s->bbox_min[c]) * dil;
        a.bbox_min[c]   = s->bbox_min[c];
        a.bbox_max[c]   = s->bbox_max[c];
    }
    a.thr_logit = 0.f;
    a.max_steps = kDefaultMaxSteps;
    if (!s->has_net) return;
    const auto& cfg = s->cfg;
    a.n_levels        = cfg.grid.n_levels;
    a.n_features      = cfg.grid.n_features_per_level;
    a.n_points        = cfg.n_points;
    a.n_inputs        = cfg.n_points * cfg.grid.n_levels * cfg.grid.n_features_per_level;
    a.n_neurons       = cfg.n_neurons;
    a.n_hidden_layers = cfg.n_hidden_layers;
    a.n_outputs       = cfg.n_outputs;
    a.depth_mode        = cfg.depth_mode;
    a.depth_linear      = cfg.depth_linear_activation;
    a.interpolate_depth = cfg.interpolate_depth;
    a.interp_eps        = cfg.depth_interpolation_eps > 0.f ? cfg.depth_interpolation_eps : 1e-3f;
    const uint32_t P  = cfg.n_points;
    float delta, offset;
    if (cfg.segment_distribution == NBVH_SEGMENT_EQUIDISTANT_EXTERIOR) {
        delta  = (P == 1) ? 1.f : 1.f / (float)(P - 1);
        offset = 0.f;
    } else {
        delta  = 1.f / (float)P;
        offset = 0.5f;
    }
    for (uint32_t k = 0; k < 4; k++) a.seg_t[k] = k < P ? delta * ((float)k + offset) : 0.f;
    a.seg_uniform = cfg.segment_distribution == NBVH_SEGMENT_UNIFORM ? 1u : 0u;
    a.seg_seed    = 0u;
    const auto lv = grid_levels(cfg.grid);
    for (size_t l = 0; l < lv.size(); l++) a.level[l] = lv[l];
    a.n_dense = 0;
    while (a.n_dense < lv.size() && lv[a.n_dense].dense) a.n_dense++;
    for (size_t l = a.n_dense; l < lv.size(); l++)
        if (lv[l].dense) a.n_dense = (uint32_t)lv.size();  // not a dense prefix: every level takes the general path
    a.tables_f32      = s->d_tables32.as<float>();
    a.tables_f16      = s->d_tables16.as<uint16_t>();
    // The MFMA kernels address the FP16 tables through a buffer resource (31-bit byte range);
    // larger tables leave tables_f16_bytes at 0 and those kernels report UNSUPPORTED.
    a.tables_f16_bytes = s->d_tables16.bytes < (1ull << 31) ? (uint32_t)s->d_tables16.bytes : 0u;
    a.table_scale     = s->table_scale;
    a.inv_table_scale = 1.f / s->table_scale;
    a.mlp_f32         = s->d_mlp32.as<float>();
    a.mlp_frag        = s->d_frag.as<uint16_t>();
    a.mlp_frag_fp8    = s->d_frag8.as<uint8_t>();
    for (int i = 0; i < 8; i++) a.fp8_wscale[i] = s->fp8_wscale[i];
}

void upload_network(nbvh_scene* s)
{
    nbvh_ctx* ctx = s->ctx;
    s->d_tables32.upload(ctx, s->tables.data(), s->tables.size());
    // FP16 tables with a power-of-two scale that brings the largest magnitude to [0.5, 1): the
    // tiny seeded values (+-1e-4) would otherwise sit next to the FP16 subnormal range. The MLP
    // is positively homogeneous (no bias, ReLU), so the scale is undone exactly on the logits.
    float amax = 0.f;
    for (float v : s->tables) amax = std::max(amax, std::fabs(v));
    int e = 0;
    if (amax > 0.f && std::isfinite(amax)) {
        std::frexp(amax, &e);  // amax = m * 2^e, m in [0.5, 1)
        e = -e;
        e = std::max(-24, std::min(24, e));
    }
    s->table_scale = std::ldexp(1.f, e);
    std::vector<uint16_t> t16(s->tables.size());
    for (size_t i = 0; i < t16.size(); i++) t16[i] = float_to_half(s->tables[i] * s->table_scale);
    s->d_tables16.upload(ctx, t16.data(), t16.size());
    const auto k32 = fp32_kernel_layout(s->cfg, s->mlp);
    s->d_mlp32.upload(ctx, k32.data(), k32.size());
    MfmaWeights mw = mfma_weight_layout(s->cfg, s->mlp);
    s->d_frag.upload(ctx, mw.frag16.data(), mw.frag16.size());
    s->d_frag8.upload(ctx, mw.frag8.data(), mw.frag8.size());
    for (int i = 0; i < 8; i++) s->fp8_wscale[i] = mw.fp8_wscale[i];
}

SlotLease::SlotLease(nbvh_ctx* ctx, hipStream_t stream) : ctx_(ctx)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(stream, &cs), "hipStreamIsCapturing");
    if (cs == hipStreamCaptureStatusActive) {
        const uint32_t c = ctx->captured_next.fetch_add(1);
        if (c >= nbvh_ctx::kCapturedSlots) {
            ctx->captured_next.store(nbvh_ctx::kCapturedSlots);
            fail(NBVH_ERR_CAPACITY, "more than 1024 queries recorded into HIP graphs on this context");
        }
        k_        = nbvh_ctx::kEagerSlots + c;
        captured_ = true;
        return;
    }
    if (cs != hipStreamCaptureStatusNone) fail(NBVH_ERR_STATE, "the stream's graph capture was invalidated");
    k_    = ctx->eager_next.fetch_add(1) % nbvh_ctx::kEagerSlots;
    lock_ = std::unique_lock<std::mutex>(ctx->slot_mu[k_]);
    hipEvent_t& ev = ctx->slot_event[k_];
    if (ev == nullptr)
        hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreateWithFlags");
    else
        hip_check(hipEventSynchronize(ev), "hipEventSynchronize");  // the slot's previous query is done
}

uint32_t* SlotLease::counters() const { return ctx_->d_counters + (size_t)k_ * nbvh_ctx::kCounterWords; }
BatchPart* SlotLease::host_parts() const { return ctx_->h_parts + (size_t)k_ * kMaxBatchParts; }
BatchPart* SlotLease::device_parts() const { return ctx_->d_parts + (size_t)k_ * kMaxBatchParts; }

void SlotLease::finish(hipStream_t stream)
{
    if (!captured_) hip_check(hipEventRecord(ctx_->slot_event[k_], stream), "hipEventRecord");
}

void ensure_part_tables(nbvh_ctx* ctx)
{
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->d_parts) return;
    const size_t bytes = (size_t)nbvh_ctx::kSlots * kMaxBatchParts * sizeof(BatchPart);
    hip_check(hipMalloc((void**)&ctx->d_parts, bytes), "hipMalloc part tables");
    if (hipHostMalloc((void**)&ctx->h_parts, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(ctx->d_parts);
        ctx->d_parts = nullptr;
        fail(NBVH_ERR_OUT_OF_MEMORY, "hipHostMalloc part tables failed");
    }
}

void run_query_ordered(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
                       hipStream_t stream, SlotLease* lease, uint32_t* fixed_counters = nullptr);
void run_query_wavefront(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
                         hipStream_t stream);

// Kernel arguments of a neural query: the scene's network plus the query parameters.
NetArgs neural_args(const nbvh_scene* s, const nbvh_query_params* p)
{
    require(p->kind <= NBVH_QUERY_TRIANGLE_ANY, "unknown query kind");
    if (s->nbvh_nodes.empty()) fail(NBVH_ERR_STATE, "neural query on a scene without an N-BVH");
    if (!s->has_net) fail(NBVH_ERR_STATE, "neural query on a scene without network weights");
    require(p->lod < NBVH_NUM_LODS, "lod must be 0..6");
    require(p->precision <= NBVH_PREC_FP8, "unknown precision");
    NetArgs a = s->args;
    a.lod       = p->lod;
    a.max_steps = p->max_traversal_steps ? p->max_traversal_steps : kDefaultMaxSteps;
    a.any_hit   = p->kind == NBVH_QUERY_NEURAL_ANY ? 1u : 0u;
    const float thr = p->visibility_threshold == 0.f ? 0.5f : p->visibility_threshold;
    require(thr > 0.f && thr < 1.f, "visibility_threshold must be in (0, 1)");
    a.thr_logit = thr == 0.5f ? 0.f : (float)std::log((double)thr / (1.0 - (double)thr));
    a.seg_seed  = p->segment_seed;
    if (a.seg_uniform && (p->flags & NBVH_QUERY_WAVEFRONT) != 0u)
        fail(NBVH_ERR_UNSUPPORTED, "the uniform segment distribution is not available in wavefront rounds");
    return a;
}

size_t query_scratch_bytes(uint32_t n)
{
    const size_t sort = std::max(ray_order_scratch_bytes(n), ray_bucket_scratch_bytes(n));
    return (sort + 255) / 256 * 256 + (size_t)n * sizeof(uint32_t);
}

constexpr uint32_t kQueryFlags = NBVH_QUERY_SORT_RAYS | NBVH_QUERY_WAVEFRONT | NBVH_QUERY_ARRAY_ORDER;
// Batches of at least this many rays with no image-order hint are traced in the coherence order
// when the scene's launch shape asks for it (ray_order 1, set by the autotuner where it was
// faster: 4M incoherent rays 28.6 ms in array order, 20.7 ms sorted; 1M camera rays are slower
// sorted) and NBVH_QUERY_ARRAY_ORDER is not set.
constexpr uint32_t kAutoSortRays = 1u << 16;

bool auto_sort_applies(const nbvh_query_params* p, uint32_t n)
{
    return (p->flags & (NBVH_QUERY_SORT_RAYS | NBVH_QUERY_ARRAY_ORDER)) == 0u && p->image_width == 0 &&
           n >= kAutoSortRays;
}

void run_query(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
               hipStream_t stream, SlotLease* lease, void* scratch)
{
    require(p != nullptr, "query params are NULL");
    require((p->flags & ~kQueryFlags) == 0u, "unknown query flags");
    if ((p->flags & NBVH_QUERY_WAVEFRONT) != 0u &&
        (p->kind == NBVH_QUERY_NEURAL_CLOSEST || p->kind == NBVH_QUERY_NEURAL_ANY)) {
        if (batch.live) fail(NBVH_ERR_UNSUPPORTED, "wavefront rounds need the ray count on the host");
        run_query_wavefront(ctx, s, p, batch, stream);
        return;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(stream, &cs), "hipStreamIsCapturing");
    bool sort = (p->flags & NBVH_QUERY_SORT_RAYS) != 0u;
    if (!sort && p->precision <= NBVH_PREC_FP8 && s->shape[p->precision].ray_order == 1u &&
        auto_sort_applies(p, batch.n) && batch.live == nullptr && cs == hipStreamCaptureStatusNone)
        sort = true;
    if (!sort || batch.n < 2 || batch.order != nullptr) {
        run_query_ordered(ctx, s, p, batch, stream, lease);
        return;
    }
    // Coherence order into the caller's scratch or transient stream-ordered memory, freed after
    // the launch: keys + radix sort; or, for a device-side count or under graph capture (the
    // radix sort records memset nodes), the bucket order of kernels alone.
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const bool buckets = batch.live != nullptr || cs != hipStreamCaptureStatusNone;
    if (scratch == nullptr && cs != hipStreamCaptureStatusNone)
        fail(NBVH_ERR_STATE, "a captured sorted query needs caller scratch (sort the rays outside the capture)");
    const size_t sort_bytes = (std::max(ray_order_scratch_bytes(batch.n), ray_bucket_scratch_bytes(batch.n)) + 255) /
                              256 * 256;
    void* mem = scratch;
    if (mem == nullptr) hip_check(hipMallocAsync(&mem, query_scratch_bytes(batch.n), stream), "hipMallocAsync");
    uint32_t* order = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + sort_bytes);
    const float* bmin = s->bbox_min;
    float ext[3];
    for (int c = 0; c < 3; c++) ext[c] = s->bbox_max[c] - s->bbox_min[c];
    try {
        if (buckets)
            hip_check(ray_bucket_order(batch, bmin, ext, order, mem, stream), "ray bucket order");
        else
            hip_check(ray_order(batch, bmin, ext, order, mem, sort_bytes, stream), "ray order");
        RayBatch sorted = batch;
        sorted.order    = order;
        run_query_ordered(ctx, s, p, sorted, stream, lease);
    } catch (...) {
        if (scratch == nullptr) (void)hipFreeAsync(mem, stream);
        throw;
    }
    if (scratch == nullptr) hip_check(hipFreeAsync(mem, stream), "hipFreeAsync");
}

// Sorted wavefront rounds (NBVH_QUERY_WAVEFRONT): see kernels/wavefront.hip. Transient
// device memory comes from the stream-ordered allocator; two small counters are read back per round.
void run_query_wavefront(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
                         hipStream_t stream)
{
    const uint32_t n = batch.n;
    if (n == 0) return;
    require(batch.parts != nullptr || (batch.rays && batch.hits), "ray or hit buffer is NULL");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    NetArgs a = neural_args(s, p);
    const uint32_t in_floats = 3 * a.n_points;
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t sort_tmp = wf_sort_scratch_bytes(n);
    size_t off = 0;
    const size_t o_st = off;      off += al((size_t)n * sizeof(WfRayState));
    const size_t o_stack = off;   off += al((size_t)n * kNeuralStackSize * sizeof(uint32_t));
    const size_t o_act0 = off;    off += al((size_t)n * 4);
    const size_t o_act1 = off;    off += al((size_t)n * 4);
    const size_t o_segs = off;    off += al((size_t)n * sizeof(WfSeg));
    const size_t o_keys = off;    off += al((size_t)n * 4);
    const size_t o_vals = off;    off += al((size_t)n * 4);
    const size_t o_kout = off;    off += al((size_t)n * 4);
    const size_t o_order = off;   off += al((size_t)n * 4);
    const size_t o_in = off;      off += al((size_t)n * in_floats * 4);
    const size_t o_logits = off;  off += al((size_t)n * 8 * 4);
    const size_t o_cnt = off;     off += 256;
    const size_t o_tmp = off;     off += al(sort_tmp);
    char* mem = nullptr;
    hip_check(hipMallocAsync((void**)&mem, off, stream), "hipMallocAsync");
    auto U32 = [&](size_t o) { return reinterpret_cast<uint32_t*>(mem + o); };
    WfRayState* st = reinterpret_cast<WfRayState*>(mem + o_st);
    WfSeg* segs    = reinterpret_cast<WfSeg*>(mem + o_segs);
    float* in      = reinterpret_cast<float*>(mem + o_in);
    float* logits  = reinterpret_cast<float*>(mem + o_logits);
    uint32_t* cnt  = U32(o_cnt);
    uint32_t* active[2] = {U32(o_act0), U32(o_act1)};
    const int end_bit = 30;  // Morton keys of segment midpoints
    auto read_u32 = [&](const uint32_t* d) {
        uint32_t v = 0;
        hip_check(hipMemcpyAsync(&v, d, 4, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return v;
    };
    try {
        hip_check(wf_init(batch, st, U32(o_stack), active[0], stream), "wavefront init");
        uint32_t n_active = n;
        int cur = 0;
        while (n_active > 0) {
            hip_check(zero_u32(cnt, 2, stream), "counter reset");
            hip_check(wf_traverse(a, batch, st, U32(o_stack), active[cur], n_active, segs, U32(o_keys), U32(o_vals),
                                  cnt, stream),
                      "wavefront traverse");
            const uint32_t m = read_u32(cnt);
            if (m == 0) break;
            hip_check(wf_sort(mem + o_tmp, al(sort_tmp), U32(o_keys), U32(o_kout), U32(o_vals), U32(o_order), m,
                              end_bit, stream),
                      "wavefront sort");
            hip_check(wf_encode(a, batch, segs, U32(o_order), m, in, stream), "wavefront encode");
            bool supported = false;
            if (p->precision == NBVH_PREC_FP32)
                hip_check(network_fp32(a, in, m, nullptr, logits, stream, &supported), "fp32 network");
            else
                hip_check(network_mfma(a, p->precision, in, m, nullptr, logits, (uint32_t)ctx->n_cus, stream, &supported), "mfma network");
            if (!supported) fail(NBVH_ERR_UNSUPPORTED, "no network kernel for this network shape / precision");
            hip_check(wf_consume(a, batch, segs, U32(o_order), logits, m, st, active[1 - cur], cnt + 1, stream),
                      "wavefront consume");
            n_active = read_u32(cnt + 1);
            cur      = 1 - cur;
        }
    } catch (...) {
        (void)hipFreeAsync(mem, stream);
        throw;
    }
    hip_check(hipFreeAsync(mem, stream), "hipFreeAsync");
}

// fixed_counters: the query's counter words are there (the allocation-free entry point) and no
// slot is taken; otherwise they come from `lease` (or a slot taken here when it is NULL).
void run_query_ordered(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
                       hipStream_t stream, SlotLease* lease, uint32_t* fixed_counters)
{
    require(p != nullptr, "query params are NULL");
    const uint32_t n = batch.n;
    if (n == 0) return;
    require(batch.parts != nullptr || (batch.rays && batch.hits), "ray or hit buffer is NULL");
    require(batch.parts == nullptr || (batch.n_parts >= 1 && batch.n_parts <= kMaxBatchParts),
            "batch part count out of range");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (p->kind == NBVH_QUERY_TRIANGLE_CLOSEST || p->kind == NBVH_QUERY_TRIANGLE_ANY) {
        if (s->bvh.empty()) fail(NBVH_ERR_STATE, "triangle query on a scene without triangles");
        TriArgs t{s->d_verts.as<float>(), s->d_tris.as<uint32_t>(), s->d_bvh.as<nbvh_node>(), s->d_prims.as<uint32_t>(),
                  p->kind == NBVH_QUERY_TRIANGLE_ANY ? 1u : 0u};
        hip_check(query_triangles(t, batch, stream), "triangle query launch");
        return;
    }
    NetArgs a = neural_args(s, p);
    // Tiled claim order only for a whole number of 8-scanline blocks (a hint: ignored otherwise).
    const uint64_t w = p->image_width;
    a.image_width = (batch.parts == nullptr && batch.order == nullptr && batch.live == nullptr && w != 0 &&
                     w % 8 == 0 && (uint64_t)n % (8 * w) == 0)
                        ? (uint32_t)w
                        : 0u;
    a.blocks_per_cu = s->shape[p->precision].blocks_per_cu;
    a.queue_bands   = s->shape[p->precision].queue_bands;
    std::unique_ptr<SlotLease> own;
    if (lease == nullptr && fixed_counters == nullptr) {
        own.reset(new SlotLease(ctx, stream));
        lease = own.get();
    }
    uint32_t* counter = fixed_counters ? fixed_counters : lease->counters();
    hip_check(zero_u32(counter, nbvh_ctx::kCounterWords, stream), "counter reset");
    bool supported = false;
    if (p->precision == NBVH_PREC_FP32)
        hip_check(query_fp32(a, batch, counter, (uint32_t)ctx->n_cus, stream, &supported), "fp32 query");
    else
        hip_check(query_mfma(a, p->precision, batch, counter, (uint32_t)ctx->n_cus, stream, &supported),
                  "mfma query");
    if (!supported) fail(NBVH_ERR_UNSUPPORTED, "no query kernel for this network shape / precision");
    if (own) own->finish(stream);
}

namespace {

void run_network(nbvh_ctx* ctx, const nbvh_scene* s, uint32_t precision, const float* d_in, uint32_t n, float* d_out,
                 float* d_logits, hipStream_t stream)
{
    if (n == 0) return;
    require(d_in != nullptr, "input buffer is NULL");
    if (!s->has_net) fail(NBVH_ERR_STATE, "network inference on a scene without weights");
    require(precision <= NBVH_PREC_FP8, "unknown precision");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    bool supported = false;
    if (precision == NBVH_PREC_FP32)
        hip_check(network_fp32(s->args, d_in, n, d_out, d_logits, stream, &supported), "fp32 network");
    else
        hip_check(network_mfma(s->args, precision, d_in, n, d_out, d_logits, (uint32_t)ctx->n_cus, stream,
                               &supported), "mfma network");
    if (!supported) fail(NBVH_ERR_UNSUPPORTED, "no network kernel for this shape / precision");
}

}  // namespace
}  // namespace nbvh

using namespace nbvh;

// =============================================================================================
// extern "C"
// =============================================================================================

NBVH_API uint32_t nbvh_abi_version(void) { return NBVH_ABI_VERSION; }

NBVH_API const char* nbvh_status_string(int32_t s)
{
    switch (s) {
    case NBVH_OK: return "ok";
    case NBVH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case NBVH_ERR_HIP: return "HIP error";
    case NBVH_ERR_NO_DEVICE: return "no device";
    case NBVH_ERR_OUT_OF_MEMORY: return "out of memory";
    case NBVH_ERR_FORMAT: return "format error";
    case NBVH_ERR_STATE: return "invalid state";
    case NBVH_ERR_UNSUPPORTED: return "unsupported";
    case NBVH_ERR_TIMEOUT: return "timeout";
    case NBVH_ERR_CAPACITY: return "capacity";
    default: return "unknown status";
    }
}

NBVH_API const char* nbvh_last_error(const nbvh_ctx* ctx)
{
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        return ctx->last_error.c_str();
    }
    return t_last_error.c_str();
}

NBVH_API int32_t nbvh_create(const nbvh_create_info* info, nbvh_ctx** out)
{
    return guarded(nullptr, [&] {
        require(out != nullptr, "out_ctx is NULL");
        *out = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            (void)hipGetLastError();
            fail(NBVH_ERR_NO_DEVICE, "no HIP device available");
        }
        const int dev = info ? info->device : 0;
        require(dev >= 0 && dev < count, "device ordinal out of range");
        hip_check(hipSetDevice(dev), "hipSetDevice");
        std::unique_ptr<nbvh_ctx> ctx(new nbvh_ctx);
        ctx->device = dev;
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        ctx->n_cus = prop.multiProcessorCount;
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        ctx->budget = (info && info->hbm_budget_bytes) ? info->hbm_budget_bytes : (uint64_t)(0.9 * (double)free_b);
        void* p = nullptr;
        const size_t counter_bytes = (size_t)nbvh_ctx::kSlots * nbvh_ctx::kCounterWords * sizeof(uint32_t);
        hip_check(hipMalloc(&p, counter_bytes), "hipMalloc counters");
        ctx->d_counters = (uint32_t*)p;
        hip_check(hipMemset(p, 0, counter_bytes), "hipMemset counters");
        // Device address of nbvh_query_noalloc's static scratch on this device (looked up once
        // per device, with the device current).
        hip_check(noalloc_scratch_init(dev), "hipGetSymbolAddress of the no-allocation scratch");
        *out = ctx.release();
    });
}

NBVH_API int32_t nbvh_destroy(nbvh_ctx* ctx)
{
    if (!ctx) return NBVH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (hipEvent_t ev : ctx->slot_event)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_parts) (void)hipFree(ctx->d_parts);
    if (ctx->h_parts) (void)hipHostFree(ctx->h_parts);
    delete ctx;
    return NBVH_OK;
}

NBVH_API int32_t nbvh_memory_usage(const nbvh_ctx* ctx, uint64_t* used, uint64_t* budget)
{
    return guarded(ctx, [&] {
        require(ctx != nullptr, "ctx is NULL");
        if (used) *used = ctx->used.load();
        if (budget) *budget = ctx->budget;
    });
}

// ----- host builders -----

NBVH_API int32_t nbvh_mesh_icosphere(uint32_t subdivisions, uint64_t seed, float amplitude, const float center[3],
                                     float radius, float* verts, uint32_t* n_verts, uint32_t* indices,
                                     uint32_t* n_tris)
{
    return guarded(nullptr, [&] {
        require(subdivisions <= 10, "subdivisions must be <= 10");
        require(n_verts && n_tris, "count outputs are NULL");
        const uint64_t p4 = 1ull << (2 * subdivisions);
        const uint64_t nt = 20 * p4, nv = 10 * p4 + 2;
        *n_verts = (uint32_t)nv;
        *n_tris  = (uint32_t)nt;
        if (!verts && !indices) return;
        require(verts && indices, "verts and indices must both be given");
        const float zero[3] = {0, 0, 0};
        std::vector<float> v;
        std::vector<uint32_t> t;
        icosphere(subdivisions, seed, amplitude, center ? center : zero, radius, v, t);
        if (v.size() != nv * 3 || t.size() != nt * 3) fail(NBVH_ERR_INVALID_ARGUMENT, "icosphere size mismatch");
        std::memcpy(verts, v.data(), v.size() * sizeof(float));
        std::memcpy(indices, t.data(), t.size() * sizeof(uint32_t));
    });
}

NBVH_API int32_t nbvh_mesh_geosphere(uint32_t frequency, uint64_t seed, float amplitude, const float center[3],
                                     float radius, float* verts, uint32_t* n_verts, uint32_t* indices,
                                     uint32_t* n_tris)
{
    return guarded(nullptr, [&] {
        require(frequency >= 1 && frequency <= 4096, "frequency must be in 1..4096");
        require(n_verts && n_tris, "count outputs are NULL");
        const uint64_t k2 = (uint64_t)frequency * frequency;
        const uint64_t nt = 20 * k2, nv = 10 * k2 + 2;
        *n_verts = (uint32_t)nv;
        *n_tris  = (uint32_t)nt;
        if (!verts && !indices) return;
        require(verts && indices, "verts and indices must both be given");
        const float zero[3] = {0, 0, 0};
        std::vector<float> v;
        std::vector<uint32_t> t;
        geosphere(frequency, seed, amplitude, center ? center : zero, radius, v, t);
        if (v.size() != nv * 3 || t.size() != nt * 3) fail(NBVH_ERR_INVALID_ARGUMENT, "geosphere size mismatch");
        std::memcpy(verts, v.data(), v.size() * sizeof(float));
        std::memcpy(indices, t.data(), t.size() * sizeof(uint32_t));
    });
}

NBVH_API int32_t nbvh_bvh_build(const float* verts, uint32_t n_verts, const uint32_t* indices, uint32_t n_tris,
                                uint32_t max_leaf_size, nbvh_node* nodes, uint32_t node_capacity, uint32_t* n_nodes,
                                uint32_t* prim_indices)
{
    return guarded(nullptr, [&] {
        require(verts && indices && nodes && n_nodes && prim_indices, "NULL argument");
        std::vector<nbvh_node> nd;
        std::vector<uint32_t> pr;
        bvh_build(verts, n_verts, indices, n_tris, max_leaf_size, nd, pr);
        *n_nodes = (uint32_t)nd.size();
        if (nd.size() > node_capacity) fail(NBVH_ERR_CAPACITY, "node capacity too small");
        std::memcpy(nodes, nd.data(), nd.size() * sizeof(nbvh_node));
        std::memcpy(prim_indices, pr.data(), pr.size() * sizeof(uint32_t));
    });
}

NBVH_API int32_t nbvh_nbvh_cut(const nbvh_node* bvh_nodes, uint32_t n_bvh_nodes, uint32_t n_leaves, nbvh_node* out,
                               uint32_t capacity, uint32_t* n_out)
{
    return guarded(nullptr, [&] {
        require(bvh_nodes && out && n_out, "NULL argument");
        std::vector<nbvh_node> nd;
        nbvh_cut(bvh_nodes, n_bvh_nodes, n_leaves, nd);
        *n_out = (uint32_t)nd.size();
        if (nd.size() > capacity) fail(NBVH_ERR_CAPACITY, "output capacity too small");
        std::memcpy(out, nd.data(), nd.size() * sizeof(nbvh_node));
    });
}

NBVH_API int32_t nbvh_nbvh_parse(const void* bytes, size_t size, nbvh_node* out, uint32_t capacity, uint32_t* n_nodes)
{
    return guarded(nullptr, [&] {
        require(bytes && n_nodes, "NULL argument");
        std::vector<nbvh_node> nd;
        nbvh_parse(bytes, size, nd);
        *n_nodes = (uint32_t)nd.size();
        if (!out) return;
        if (nd.size() > capacity) fail(NBVH_ERR_CAPACITY, "node capacity too small");
        std::memcpy(out, nd.data(), nd.size() * sizeof(nbvh_node));
    });
}

NBVH_API int32_t nbvh_nbvh_serialize(const nbvh_node* nodes, uint32_t n_nodes, void* out, size_t capacity,
                                     size_t* size)
{
    return guarded(nullptr, [&] {
        require(nodes && size, "NULL argument");
        const auto buf = nbvh_serialize(nodes, n_nodes);
        *size = buf.size();
        if (!out) return;
        if (buf.size() > capacity) fail(NBVH_ERR_CAPACITY, "output capacity too small");
        std::memcpy(out, buf.data(), buf.size());
    });
}

NBVH_API int32_t nbvh_network_param_counts(const nbvh_network_config* cfg, uint64_t* n_table, uint64_t* n_mlp)
{
    return guarded(nullptr, [&] {
        require(cfg != nullptr, "cfg is NULL");
        network_validate(*cfg);
        if (n_table) *n_table = table_param_count(*cfg);
        if (n_mlp) *n_mlp = mlp_param_count(*cfg);
    });
}

NBVH_API int32_t nbvh_tcnn_params_to_network(const nbvh_network_config* cfg, const float* params, uint64_t n_params,
                                             float* tables, float* mlp)
{
    return guarded(nullptr, [&] {
        require(cfg && params && tables && mlp, "NULL argument");
        network_validate(*cfg);
        if (n_params != tcnn_param_count(*cfg))
            fail(NBVH_ERR_INVALID_ARGUMENT, "tcnn parameter count " + std::to_string(n_params) + " != " +
                                                std::to_string(tcnn_param_count(*cfg)) + " for this network");
        tcnn_params_split(*cfg, params, tables, mlp);
    });
}

NBVH_API int32_t nbvh_network_to_tcnn_params(const nbvh_network_config* cfg, const float* tables, const float* mlp,
                                             float* params, uint64_t capacity, uint64_t* n_params)
{
    return guarded(nullptr, [&] {
        require(cfg != nullptr, "cfg is NULL");
        network_validate(*cfg);
        const uint64_t n = tcnn_param_count(*cfg);
        if (n_params) *n_params = n;
        if (!params) return;
        require(tables && mlp, "NULL argument");
        if (capacity < n) fail(NBVH_ERR_CAPACITY, "parameter buffer too small");
        tcnn_params_join(*cfg, tables, mlp, params);
    });
}

NBVH_API int32_t nbvh_network_init_params(const nbvh_network_config* cfg, uint64_t seed, float table_range,
                                          float* tables, float* mlp)
{
    return guarded(nullptr, [&] {
        require(cfg && tables && mlp, "NULL argument");
        network_validate(*cfg);
        network_init(*cfg, seed, table_range, tables, mlp);
    });
}

NBVH_API int32_t nbvh_camera_look_at(const float eye[3], const float target[3], const float up[3], float fovy,
                                     uint32_t width, uint32_t height, nbvh_camera* out)
{
    return guarded(nullptr, [&] {
        require(eye && target && up && out, "NULL argument");
        require(width > 0 && height > 0, "image size must be positive");
        camera_look_at(eye, target, up, fovy, width, height, *out);
    });
}

// ----- scenes -----

NBVH_API int32_t nbvh_scene_create(nbvh_ctx* ctx, const float* verts, uint32_t n_verts, const uint32_t* indices,
                                   uint32_t n_tris, nbvh_scene** out)
{
    return guarded(ctx, [&] {
        require(ctx && out, "NULL argument");
        require(verts && indices && n_tris > 0, "empty mesh");
        *out = nullptr;
        std::unique_ptr<nbvh_scene> s(new nbvh_scene);
        s->ctx = ctx;
        s->verts.assign(verts, verts + (size_t)n_verts * 3);
        s->tris.assign(indices, indices + (size_t)n_tris * 3);
        bvh_build(verts, n_verts, indices, n_tris, 2, s->bvh, s->prims);
        for (int c = 0; c < 3; c++) {
            s->bbox_min[c] = s->bvh[0].aabb_min[c];
            s->bbox_max[c] = s->bvh[0].aabb_max[c];
        }
        s->has_bbox = true;
        s->d_verts.upload(ctx, s->verts.data(), s->verts.size());
        s->d_tris.upload(ctx, s->tris.data(), s->tris.size());
        s->d_bvh.upload(ctx, s->bvh.data(), s->bvh.size());
        s->d_prims.upload(ctx, s->prims.data(), s->prims.size());
        refresh_args(s.get());
        *out = s.release();
    });
}

NBVH_API int32_t nbvh_scene_create_empty(nbvh_ctx* ctx, nbvh_scene** out)
{
    return guarded(ctx, [&] {
        require(ctx && out, "NULL argument");
        std::unique_ptr<nbvh_scene> s(new nbvh_scene);
        s->ctx = ctx;
        refresh_args(s.get());
        *out = s.release();
    });
}

NBVH_API int32_t nbvh_scene_destroy(nbvh_scene* s)
{
    if (!s) return NBVH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipDeviceSynchronize();
    delete s;
    return NBVH_OK;
}

NBVH_API int32_t nbvh_scene_get_bvh(const nbvh_scene* s, nbvh_node* nodes, uint32_t node_capacity, uint32_t* n_nodes,
                                    uint32_t* prim_indices, uint32_t prim_capacity, uint32_t* n_prims)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s != nullptr, "scene is NULL");
        if (n_nodes) *n_nodes = (uint32_t)s->bvh.size();
        if (n_prims) *n_prims = (uint32_t)s->prims.size();
        if (nodes) {
            if (s->bvh.size() > node_capacity) fail(NBVH_ERR_CAPACITY, "node capacity too small");
            std::memcpy(nodes, s->bvh.data(), s->bvh.size() * sizeof(nbvh_node));
        }
        if (prim_indices) {
            if (s->prims.size() > prim_capacity) fail(NBVH_ERR_CAPACITY, "prim capacity too small");
            std::memcpy(prim_indices, s->prims.data(), s->prims.size() * sizeof(uint32_t));
        }
    });
}

void nbvh::set_scene_nbvh(nbvh_scene* s, std::vector<nbvh_node> nodes)
{
    nbvh_validate(nodes.data(), (uint32_t)nodes.size());
    const auto dil = nbvh_dilate(nodes, kFixedNodeDilation);
    s->d_nbvh.upload(s->ctx, dil.data(), dil.size());
    s->nbvh_nodes = std::move(nodes);
    if (!s->has_bbox) {
        // Without a mesh the encoding domain is the N-BVH root box (the reference uses the neural
        // mesh's BVH root box, neural_bvh_renderer_accel.cpp:248-249, which equals it for a cut).
        for (int c = 0; c < 3; c++) {
            s->bbox_min[c] = s->nbvh_nodes[0].aabb_min[c];
            s->bbox_max[c] = s->nbvh_nodes[0].aabb_max[c];
        }
        s->has_bbox = true;
    }
    refresh_args(s);
}

NBVH_API int32_t nbvh_scene_build_nbvh(nbvh_scene* s, uint32_t n_leaves)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s != nullptr, "scene is NULL");
        if (s->bvh.empty()) fail(NBVH_ERR_STATE, "scene has no triangle BVH to cut");
        std::vector<nbvh_node> nd;
        nbvh_cut(s->bvh.data(), (uint32_t)s->bvh.size(), n_leaves, nd);
        set_scene_nbvh(s, std::move(nd));
    });
}

NBVH_API int32_t nbvh_scene_load_nbvh(nbvh_scene* s, const void* bytes, size_t size)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s != nullptr, "scene is NULL");
        std::vector<nbvh_node> nd;
        nbvh_parse(bytes, size, nd);
        set_scene_nbvh(s, std::move(nd));
    });
}

NBVH_API int32_t nbvh_scene_set_nbvh(nbvh_scene* s, const nbvh_node* nodes, uint32_t n_nodes)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && nodes && n_nodes > 0, "NULL argument");
        set_scene_nbvh(s, std::vector<nbvh_node>(nodes, nodes + n_nodes));
    });
}

NBVH_API int32_t nbvh_scene_get_nbvh(const nbvh_scene* s, nbvh_node* nodes, uint32_t capacity, uint32_t* n_nodes)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && n_nodes, "NULL argument");
        *n_nodes = (uint32_t)s->nbvh_nodes.size();
        if (!nodes) return;
        if (s->nbvh_nodes.size() > capacity) fail(NBVH_ERR_CAPACITY, "node capacity too small");
        std::memcpy(nodes, s->nbvh_nodes.data(), s->nbvh_nodes.size() * sizeof(nbvh_node));
    });
}

NBVH_API int32_t nbvh_scene_set_bbox(nbvh_scene* s, const float bmin[3], const float bmax[3])
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && bmin && bmax, "NULL argument");
        for (int c = 0; c < 3; c++) require(bmin[c] < bmax[c], "bbox must have positive extent");
        std::memcpy(s->bbox_min, bmin, sizeof(s->bbox_min));
        std::memcpy(s->bbox_max, bmax, sizeof(s->bbox_max));
        s->has_bbox = true;
        refresh_args(s);
    });
}

NBVH_API int32_t nbvh_scene_get_bbox(const nbvh_scene* s, float bmin[3], float bmax[3])
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && bmin && bmax, "NULL argument");
        std::memcpy(bmin, s->bbox_min, sizeof(s->bbox_min));
        std::memcpy(bmax, s->bbox_max, sizeof(s->bbox_max));
    });
}

NBVH_API int32_t nbvh_scene_load_network(nbvh_scene* s, const nbvh_network_config* cfg, const float* tables,
                                         uint64_t n_table, const float* mlp, uint64_t n_mlp)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && cfg && tables && mlp, "NULL argument");
        network_validate(*cfg);
        if (n_table != table_param_count(*cfg)) fail(NBVH_ERR_FORMAT, "table parameter count does not match config");
        if (n_mlp != mlp_param_count(*cfg)) fail(NBVH_ERR_FORMAT, "MLP parameter count does not match config");
        for (uint64_t i = 0; i < n_mlp; i++)
            if (!std::isfinite(mlp[i])) fail(NBVH_ERR_FORMAT, "MLP weights contain NaN/inf");
        s->cfg = *cfg;
        s->tables.assign(tables, tables + n_table);
        s->mlp.assign(mlp, mlp + n_mlp);
        upload_network(s);
        s->has_net = true;
        s->net_generation++;
        refresh_args(s);
    });
}

NBVH_API int32_t nbvh_scene_init_network(nbvh_scene* s, const nbvh_network_config* cfg, uint64_t seed,
                                         float table_range)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && cfg, "NULL argument");
        network_validate(*cfg);
        std::vector<float> t(table_param_count(*cfg)), m(mlp_param_count(*cfg));
        network_init(*cfg, seed, table_range, t.data(), m.data());
        s->cfg = *cfg;
        s->tables.swap(t);
        s->mlp.swap(m);
        upload_network(s);
        s->has_net = true;
        s->net_generation++;
        refresh_args(s);
    });
}

NBVH_API int32_t nbvh_scene_get_network(const nbvh_scene* s, nbvh_network_config* cfg, float* tables,
                                        uint64_t table_capacity, float* mlp, uint64_t mlp_capacity)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s != nullptr, "scene is NULL");
        if (!s->has_net) fail(NBVH_ERR_STATE, "scene has no network");
        if (cfg) *cfg = s->cfg;
        if (tables) {
            if (s->tables.size() > table_capacity) fail(NBVH_ERR_CAPACITY, "table capacity too small");
            std::memcpy(tables, s->tables.data(), s->tables.size() * sizeof(float));
        }
        if (mlp) {
            if (s->mlp.size() > mlp_capacity) fail(NBVH_ERR_CAPACITY, "MLP capacity too small");
            std::memcpy(mlp, s->mlp.data(), s->mlp.size() * sizeof(float));
        }
    });
}

// ----- queries -----

NBVH_API int32_t nbvh_raygen(nbvh_ctx* ctx, const nbvh_camera* cam, uint32_t w, uint32_t h, uint32_t sample,
                             uint32_t frame, nbvh_ray* d_rays, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && cam && d_rays, "NULL argument");
        require((uint64_t)w * h <= 0xFFFFFFFFull, "image too large");
        if ((uint64_t)w * h == 0) return;
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hip_check(raygen(*cam, w, h, sample, frame, d_rays, (hipStream_t)stream), "raygen launch");
    });
}

NBVH_API int32_t nbvh_scatter_diffuse(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_ray* d_rays,
                                      const nbvh_hit* d_hits, uint32_t n, uint32_t sample, uint32_t bounce,
                                      nbvh_ray* d_next, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        if (n == 0) return;
        require(d_rays && d_hits && d_next, "NULL buffer");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        TriArgs t{s->d_verts.as<float>(), s->d_tris.as<uint32_t>(), s->d_bvh.as<nbvh_node>(), s->d_prims.as<uint32_t>(),
                  0u};
        hip_check(scatter_diffuse(t, d_rays, d_hits, n, sample, bounce, d_next, (hipStream_t)stream), "scatter launch");
    });
}

NBVH_API int32_t nbvh_segment_labels(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_ray* d_rays, const float* d_mint,
                                     const float* d_maxt, uint32_t n, float* d_targets, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        if (n == 0) return;
        require(d_rays && d_mint && d_maxt && d_targets, "NULL buffer");
        if (s->bvh.empty()) fail(NBVH_ERR_STATE, "segment labels need the scene's triangles");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        TriArgs t{s->d_verts.as<float>(), s->d_tris.as<uint32_t>(), s->d_bvh.as<nbvh_node>(), s->d_prims.as<uint32_t>(),
                  0u};
        hip_check(segment_labels(t, d_rays, d_mint, d_maxt, n, d_targets, (hipStream_t)stream), "labels launch");
    });
}

NBVH_API int32_t nbvh_query(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const nbvh_ray* d_rays,
                            uint32_t n, nbvh_hit* d_hits, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        run_query(ctx, s, p, RayBatch{d_rays, d_hits, nullptr, 0, n}, (hipStream_t)stream);
    });
}

NBVH_API int32_t nbvh_query_noalloc(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p,
                                    const nbvh_ray* d_rays, uint32_t n, nbvh_hit* d_hits, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s && p, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        require((p->flags & ~kQueryFlags) == 0u, "unknown query flags");
        if ((p->flags & NBVH_QUERY_WAVEFRONT) != 0u)
            fail(NBVH_ERR_UNSUPPORTED, "nbvh_query_noalloc: the wavefront-round schedule needs allocated scratch");
        if (n == 0) return;
        require(d_rays && d_hits, "ray or hit buffer is NULL");
        const bool sort = (p->flags & NBVH_QUERY_SORT_RAYS) != 0u && n >= 2;
        if (sort && n > NBVH_NOALLOC_MAX_SORTED_RAYS)
            fail(NBVH_ERR_CAPACITY, "nbvh_query_noalloc: a sorted query takes at most " +
                                        std::to_string(NBVH_NOALLOC_MAX_SORTED_RAYS) + " rays, got " + std::to_string(n));
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        char* base = static_cast<char*>(noalloc_scratch(ctx->device));
        if (base == nullptr) fail(NBVH_ERR_STATE, "no-allocation scratch not set up on this device");
        // Layout: [0, 256) the counter words; [256, 256 + sort bytes) keys, ids and the radix
        // sort's storage; then the claim order (n words).
        constexpr size_t kCounterBytes = 256;
        static_assert(nbvh_ctx::kCounterWords * sizeof(uint32_t) <= kCounterBytes, "counter words exceed their area");
        uint32_t* counters = reinterpret_cast<uint32_t*>(base);
        const hipStream_t st = (hipStream_t)stream;
        RayBatch batch{d_rays, d_hits, nullptr, 0, n};
        if (!sort) {
            run_query_ordered(ctx, s, p, batch, st, nullptr, counters);
            return;
        }
        const size_t sort_bytes = (ray_order_scratch_bytes(n) + 255) / 256 * 256;
        const size_t need = kCounterBytes + sort_bytes + (size_t)n * sizeof(uint32_t);
        if (need > kNoAllocScratchBytes)
            fail(NBVH_ERR_CAPACITY, "nbvh_query_noalloc: sorting " + std::to_string(n) + " rays needs " +
                                        std::to_string(need) + " B of scratch, more than its " +
                                        std::to_string(kNoAllocScratchBytes) + " B");
        void* sort_mem = base + kCounterBytes;
        uint32_t* order = reinterpret_cast<uint32_t*>(base + kCounterBytes + sort_bytes);
        float ext[3];
        for (int c = 0; c < 3; c++) ext[c] = s->bbox_max[c] - s->bbox_min[c];
        hip_check(ray_order(batch, s->bbox_min, ext, order, sort_mem, sort_bytes, st), "ray order");
        RayBatch sorted = batch;
        sorted.order    = order;
        run_query_ordered(ctx, s, p, sorted, st, nullptr, counters);
    });
}

namespace nbvh {
namespace {

// Queue one multi-part query (part table into a scratch slot, one launch).
void run_parts(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const nbvh_query_part* parts,
               uint32_t n_parts, hipStream_t stream)
{
    ensure_part_tables(ctx);
    SlotLease lease(ctx, stream);
    BatchPart* h = lease.host_parts();
    BatchPart* d = lease.device_parts();
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_parts; i++) {
        require(parts[i].n_rays == 0 || (parts[i].rays && parts[i].hits), "NULL ray or hit buffer in a part");
        require(parts[i].any_hit <= 1u, "any_hit must be 0 or 1");
        h[i] = BatchPart{parts[i].rays, parts[i].hits, (uint32_t)total, parts[i].n_rays, parts[i].any_hit, 0u};
        total += parts[i].n_rays;
    }
    require(total < (1ull << 32), "more than 2^32 - 1 rays in one query");
    if (total == 0) return;
    hip_check(hipMemcpyAsync(d, h, n_parts * sizeof(BatchPart), hipMemcpyHostToDevice, stream), "part table H2D");
    run_query(ctx, s, p, RayBatch{nullptr, nullptr, d, n_parts, (uint32_t)total}, stream, &lease);
    lease.finish(stream);
}

}  // namespace
}  // namespace nbvh

NBVH_API int32_t nbvh_query_parts(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p,
                                  const nbvh_query_part* parts, uint32_t n_parts, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s && p && parts, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        require(n_parts >= 1 && n_parts <= kMaxBatchParts, "n_parts must be in 1..64");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        run_parts(ctx, s, p, parts, n_parts, (hipStream_t)stream);
    });
}

NBVH_API int32_t nbvh_scene_set_launch_shape(nbvh_scene* s, uint32_t precision, const nbvh_launch_shape* shape)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && shape, "NULL argument");
        require(precision <= NBVH_PREC_FP8, "unknown precision");
        require(shape->queue_bands <= 8, "queue_bands must be 0..8");
        require(shape->ray_order <= 1, "ray_order must be 0 or 1");
        s->shape[precision] = *shape;
    });
}

NBVH_API int32_t nbvh_scene_get_launch_shape(const nbvh_scene* s, uint32_t precision, nbvh_launch_shape* shape)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && shape, "NULL argument");
        require(precision <= NBVH_PREC_FP8, "unknown precision");
        *shape = s->shape[precision];
    });
}

namespace nbvh {
namespace {

// Launch-shape autotuner: persistent workgroups per CU x ray-queue bands, each candidate timed
// with events over kReps launches of `run` (one query of n rays under the scene's current shape);
// then the winner in the coherence order where that applies to the batch. The fastest is kept
// for the scene and precision. The shape only changes which wave traces which ray, never a result.
template <typename Run>
void autotune_batch(nbvh_scene* s, const nbvh_query_params* p, uint32_t n, hipStream_t st, const Run& run,
                    nbvh_launch_shape* best, float* best_ms)
{
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        fail(NBVH_ERR_HIP, "hipEventCreate failed");
    }
    const nbvh_launch_shape saved = s->shape[p->precision];
    nbvh_launch_shape win = saved;
    float win_ms = FLT_MAX;
    constexpr int kReps = 3;
    auto time_shape = [&](const nbvh_query_params* q) {
        run(q);  // warm-up
        hip_check(hipEventRecord(e0, st), "hipEventRecord");
        for (int r = 0; r < kReps; r++) run(q);
        hip_check(hipEventRecord(e1, st), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / kReps;
    };
    try {
        // Shapes in array order (unless the caller asked for the coherence order), then the
        // winner in the coherence order where that would apply by default.
        nbvh_query_params arr = *p;
        if ((arr.flags & NBVH_QUERY_SORT_RAYS) == 0u) arr.flags |= NBVH_QUERY_ARRAY_ORDER;
        for (uint32_t bpc : {0u, 1u, 2u, 3u})
            for (uint32_t bands : {1u, 2u, 4u, 8u}) {
                s->shape[p->precision] = nbvh_launch_shape{bpc, bands, 0u, 0u};
                const float ms = time_shape(&arr);
                if (ms < win_ms) {
                    win_ms = ms;
                    win    = s->shape[p->precision];
                }
            }
        if (auto_sort_applies(p, n)) {
            nbvh_query_params srt = *p;
            srt.flags |= NBVH_QUERY_SORT_RAYS;
            s->shape[p->precision] = win;
            const float ms = time_shape(&srt);
            if (ms < win_ms) {
                win_ms        = ms;
                win.ray_order = 1u;
            }
        }
    } catch (...) {
        s->shape[p->precision] = saved;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    s->shape[p->precision] = win;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
}

}  // namespace
}  // namespace nbvh

NBVH_API int32_t nbvh_autotune(nbvh_ctx* ctx, nbvh_scene* s, const nbvh_query_params* p, const nbvh_ray* d_rays,
                               uint32_t n, nbvh_hit* d_hits, void* stream, nbvh_launch_shape* best, float* best_ms)
{
    return guarded(ctx, [&] {
        require(ctx && s && p, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        require(p->precision <= NBVH_PREC_FP8, "unknown precision");
        require(n > 0 && d_rays && d_hits, "empty ray batch");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = (hipStream_t)stream;
        const RayBatch batch{d_rays, d_hits, nullptr, 0, n};
        autotune_batch(s, p, n, st, [&](const nbvh_query_params* q) { run_query(ctx, s, q, batch, st); }, best,
                       best_ms);
    });
}

NBVH_API int32_t nbvh_autotune_parts(nbvh_ctx* ctx, nbvh_scene* s, const nbvh_query_params* p,
                                     const nbvh_query_part* parts, uint32_t n_parts, void* stream,
                                     nbvh_launch_shape* best, float* best_ms)
{
    return guarded(ctx, [&] {
        require(ctx && s && p && parts, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        require(p->precision <= NBVH_PREC_FP8, "unknown precision");
        require(n_parts >= 1 && n_parts <= kMaxBatchParts, "n_parts must be in 1..64");
        uint64_t total = 0;
        for (uint32_t i = 0; i < n_parts; i++) total += parts[i].n_rays;
        require(total > 0 && total < (1ull << 32), "the parts hold no rays or more than 2^32 - 1");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = (hipStream_t)stream;
        autotune_batch(s, p, (uint32_t)total, st,
                       [&](const nbvh_query_params* q) { run_parts(ctx, s, q, parts, n_parts, st); }, best, best_ms);
    });
}

namespace nbvh {

size_t hybrid_scratch_bytes(uint32_t n, bool sort)
{
    const size_t bufs = ((size_t)n * (sizeof(nbvh_ray) + sizeof(nbvh_hit)) + 255) / 256 * 256;
    return bufs + (sort ? query_scratch_bytes(n) : 0);
}

// Hybrid closest/any hit (see nbvh_query_hybrid): triangles first, the neural query in front of
// each triangle hit, the neural record replacing the triangle's when there is one.
void run_hybrid(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* triangles, const nbvh_query_params* p,
                const nbvh_ray* d_rays, uint32_t n, nbvh_hit* d_hits, hipStream_t st, const uint32_t* live,
                void* scratch)
{
    const bool any = p->kind == NBVH_QUERY_NEURAL_ANY;
    nbvh_query_params tp = *p;
    tp.kind  = any ? NBVH_QUERY_TRIANGLE_ANY : NBVH_QUERY_TRIANGLE_CLOSEST;
    tp.flags = p->flags & (NBVH_QUERY_SORT_RAYS | NBVH_QUERY_ARRAY_ORDER);
    const size_t bufs = hybrid_scratch_bytes(n, false);
    void* sort_scratch = scratch ? static_cast<char*>(scratch) + bufs : nullptr;
    RayBatch tb{d_rays, d_hits, nullptr, 0, n};
    tb.live = live;
    run_query(ctx, triangles, &tp, tb, st, nullptr, sort_scratch);
    void* mem = scratch;
    if (mem == nullptr) hip_check(hipMallocAsync(&mem, bufs, st), "hipMallocAsync");
    nbvh_ray* clipped = static_cast<nbvh_ray*>(mem);
    nbvh_hit* nh      = reinterpret_cast<nbvh_hit*>(clipped + n);
    try {
        hip_check(hybrid_clip(d_rays, d_hits, n, live, clipped, st), "hybrid clip");
        RayBatch nb{clipped, nh, nullptr, 0, n};
        nb.live = live;
        run_query(ctx, neural, p, nb, st, nullptr, sort_scratch);
        hip_check(hybrid_merge(d_hits, nh, n, live, any, st), "hybrid merge");
    } catch (...) {
        if (scratch == nullptr) (void)hipFreeAsync(mem, st);
        throw;
    }
    if (scratch == nullptr) hip_check(hipFreeAsync(mem, st), "hipFreeAsync");
}

}  // namespace nbvh

NBVH_API int32_t nbvh_query_hybrid(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* triangles,
                                   const nbvh_query_params* p, const nbvh_ray* d_rays, uint32_t n, nbvh_hit* d_hits,
                                   void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && neural && triangles && p, "NULL argument");
        require(neural->ctx == ctx && triangles->ctx == ctx, "scene belongs to another context");
        require(p->kind == NBVH_QUERY_NEURAL_CLOSEST || p->kind == NBVH_QUERY_NEURAL_ANY,
                "hybrid queries take a neural kind (closest or any)");
        if (n == 0) return;
        require(d_rays && d_hits, "NULL buffer");
        if (triangles->bvh.empty()) fail(NBVH_ERR_STATE, "the triangle scene has no triangles");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        run_hybrid(ctx, neural, triangles, p, d_rays, n, d_hits, (hipStream_t)stream);
    });
}

namespace nbvh {
namespace {

// Device memory of one nbvh_path_trace frame, one block: ping-pong rays and path states, the hit
// records, the live-path counts (counts[d]: paths that reach depth d + 1) and the queries' scratch.
struct FrameLayout {
    static constexpr uint32_t kCountWords = 64;
    size_t rays[2], state[2], hits, counts, scratch, total;
    FrameLayout(uint32_t n, const nbvh_path_params* p, bool hybrid)
    {
        auto al = [](size_t b) { return (b + 255) / 256 * 256; };
        // Sort scratch: asked for, or the automatic coherence order of primary rays without an
        // image-order hint.
        const bool sort = ((p->query.flags | p->bounce_query_flags) & NBVH_QUERY_SORT_RAYS) != 0u ||
                          ((p->query.flags & NBVH_QUERY_ARRAY_ORDER) == 0u && p->query.image_width == 0 &&
                           n >= kAutoSortRays);
        size_t off = 0;
        for (int k = 0; k < 2; k++) {
            rays[k] = off;
            off += al((size_t)n * sizeof(nbvh_ray));
        }
        for (int k = 0; k < 2; k++) {
            state[k] = off;
            off += al((size_t)n * sizeof(PathState));
        }
        hits = off;
        off += al((size_t)n * sizeof(nbvh_hit));
        counts = off;
        off += kCountWords * sizeof(uint32_t);
        scratch = off;
        off += al(hybrid ? hybrid_scratch_bytes(n, sort) : sort ? query_scratch_bytes(n) : 0);
        total = off;
    }
};

void check_path_params(const nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* environment,
                       const nbvh_path_params* p, uint32_t width, uint32_t height)
{
    require(ctx && neural && p, "NULL argument");
    require(neural->ctx == ctx && (!environment || environment->ctx == ctx), "scene belongs to another context");
    const bool tri_object = p->query.kind == NBVH_QUERY_TRIANGLE_CLOSEST;
    require(p->query.kind == NBVH_QUERY_NEURAL_CLOSEST || tri_object,
            "path tracing takes NBVH_QUERY_NEURAL_CLOSEST or NBVH_QUERY_TRIANGLE_CLOSEST");
    require(!(tri_object && environment), "a triangle rendering of the object takes no environment");
    require(p->secondary_visibility_threshold == 0.f ||
                (p->secondary_visibility_threshold > 0.f && p->secondary_visibility_threshold < 1.f),
            "secondary_visibility_threshold must be 0 or in (0, 1)");
    if (tri_object && neural->bvh.empty()) fail(NBVH_ERR_STATE, "the object has no triangles");
    require(p->max_depth >= 1 && p->max_depth <= 16, "max_depth must be 1..16");
    require(p->background <= NBVH_BACKGROUND_SKY, "unknown background");
    require((uint64_t)width * height <= 0xFFFFFFFFull, "image too large");
    if (environment && environment->bvh.empty()) fail(NBVH_ERR_STATE, "the environment has no triangles");
}

}  // namespace
}  // namespace nbvh

NBVH_API int32_t nbvh_path_trace_workspace_size(nbvh_ctx* ctx, const nbvh_scene* neural,
                                                const nbvh_scene* environment, const nbvh_path_params* p,
                                                uint32_t width, uint32_t height, uint64_t* bytes)
{
    return guarded(ctx, [&] {
        require(bytes != nullptr, "bytes is NULL");
        check_path_params(ctx, neural, environment, p, width, height);
        *bytes = (uint64_t)FrameLayout(width * height, p, environment != nullptr).total;
    });
}

NBVH_API int32_t nbvh_path_trace(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* environment,
                                 const nbvh_path_params* p, const nbvh_camera* cam, uint32_t width, uint32_t height,
                                 float* d_accum, nbvh_path_stats* stats, void* stream)
{
    return guarded(ctx, [&] {
        check_path_params(ctx, neural, environment, p, width, height);
        require(cam != nullptr, "NULL argument");
        const bool tri_object = p->query.kind == NBVH_QUERY_TRIANGLE_CLOSEST;
        // A path's rays move between array positions from depth to depth, and a uniform segment
        // distribution seeds each ray's stream by its position: not reproducible in a frame.
        if (!tri_object && neural->has_net && neural->cfg.segment_distribution == NBVH_SEGMENT_UNIFORM)
            fail(NBVH_ERR_UNSUPPORTED, "path tracing a network with the uniform segment distribution");
        if (stats) *stats = nbvh_path_stats{};
        const uint32_t n = width * height;
        if (n == 0) return;
        require(d_accum != nullptr, "NULL accumulation buffer");
        hipStream_t st = (hipStream_t)stream;
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        // Wavefront rounds read their ray counts on the host: with them, the frame reads each
        // depth's live-path count back; otherwise the counts stay on the device.
        const bool host_counts = (p->bounce_query_flags & NBVH_QUERY_WAVEFRONT) != 0u;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
        const bool capturing = cs != hipStreamCaptureStatusNone;
        const FrameLayout lay(n, p, environment != nullptr);
        if (capturing) {
            require(p->workspace != nullptr, "a captured frame needs a workspace (nbvh_path_trace_workspace_size)");
            require(stats == nullptr, "stats must be NULL while the stream is captured");
            require(!host_counts, "wavefront-round bounces cannot be captured");
        }
        if (p->workspace) require(p->workspace_bytes >= lay.total, "workspace smaller than nbvh_path_trace_workspace_size");
        char* mem = static_cast<char*>(p->workspace);
        if (mem == nullptr) hip_check(hipMallocAsync((void**)&mem, lay.total, st), "hipMallocAsync");
        nbvh_ray* rays[2] = {reinterpret_cast<nbvh_ray*>(mem + lay.rays[0]), reinterpret_cast<nbvh_ray*>(mem + lay.rays[1])};
        PathState* state[2] = {reinterpret_cast<PathState*>(mem + lay.state[0]),
                               reinterpret_cast<PathState*>(mem + lay.state[1])};
        nbvh_hit* hits  = reinterpret_cast<nbvh_hit*>(mem + lay.hits);
        uint32_t* cnt   = reinterpret_cast<uint32_t*>(mem + lay.counts);
        void* scratch   = lay.total > lay.scratch ? mem + lay.scratch : nullptr;
        PathArgs a{};
        const nbvh_scene* tri_scene = tri_object ? neural : environment;
        if (tri_scene)
            a.tri = TriArgs{tri_scene->d_verts.as<float>(), tri_scene->d_tris.as<uint32_t>(),
                            tri_scene->d_bvh.as<nbvh_node>(), tri_scene->d_prims.as<uint32_t>(), 0u};
        a.accum          = d_accum;
        a.max_depth      = p->max_depth;
        a.rr_start_depth = p->rr_start_depth;
        a.sample         = p->sample;
        a.background     = p->background;
        for (int c = 0; c < 3; c++) {
            a.background_color[c] = p->background_color[c];
            a.env_albedo[c]       = p->environment_albedo[c];
        }
        a.normal_offset = p->normal_offset > 0.f ? p->normal_offset : 1e-4f;
        try {
            hip_check(zero_u32(cnt, FrameLayout::kCountWords, st), "counter reset");
            hip_check(raygen(*cam, width, height, p->sample, 0u, rays[0], st), "raygen launch");
            hip_check(path_init(state[0], n, st), "path init");
            uint32_t cur = n;  // host-side count: depth 1, and every depth with host_counts
            int c = 0;
            for (uint32_t depth = 1; depth <= p->max_depth; depth++) {
                // Depth 1 traces all n primary rays; deeper rays are the live paths counted on
                // the device (launches sized for n), or read back with host_counts.
                const uint32_t* live = depth == 1 || host_counts ? nullptr : cnt + (depth - 1);
                const uint32_t cap   = host_counts ? cur : n;
                if (cap == 0) break;
                nbvh_query_params q = p->query;
                if (depth > 1) {
                    q.flags       = p->bounce_query_flags;
                    q.image_width = 0;
                    if (p->path_lod_switch) q.lod = p->lod_switch;
                    if (p->secondary_visibility_threshold > 0.f)
                        q.visibility_threshold = p->secondary_visibility_threshold;
                }
                if (environment) {
                    run_hybrid(ctx, neural, environment, &q, rays[c], cap, hits, st, live, scratch);
                } else {
                    RayBatch rb{rays[c], hits, nullptr, 0, cap};
                    rb.live = live;
                    run_query(ctx, neural, &q, rb, st, nullptr, scratch);
                }
                hip_check(path_shade(a, depth, cap, live, rays[c], hits, state[c], rays[1 - c], state[1 - c],
                                     cnt + depth, st),
                          "path shade");
                if (depth == p->max_depth) break;
                if (host_counts) {
                    hip_check(hipMemcpyAsync(&cur, cnt + depth, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
                    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
                }
                c = 1 - c;
            }
            if (stats) {
                uint32_t h[FrameLayout::kCountWords];
                hip_check(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
                stats->rays[0] = n;
                for (uint32_t d = 2; d <= p->max_depth; d++) stats->rays[d - 1] = h[d - 1];
            }
        } catch (...) {
            if (p->workspace == nullptr) (void)hipFreeAsync(mem, st);
            throw;
        }
        if (p->workspace == nullptr) hip_check(hipFreeAsync(mem, st), "hipFreeAsync");
    });
}

NBVH_API int32_t nbvh_query_host(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const nbvh_ray* rays,
                                 uint32_t n, nbvh_hit* hits)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        require(s->ctx == ctx, "scene belongs to another context");
        if (n == 0) return;
        require(rays && hits, "NULL buffer");
        DevBuf dr, dh;
        dr.upload(ctx, rays, n);
        dh.alloc(ctx, (size_t)n * sizeof(nbvh_hit));
        run_query(ctx, s, p, RayBatch{dr.as<nbvh_ray>(), dh.as<nbvh_hit>(), nullptr, 0, n}, nullptr);
        hip_check(hipMemcpy(hits, dh.ptr, (size_t)n * sizeof(nbvh_hit), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    });
}

NBVH_API int32_t nbvh_network_inference(nbvh_ctx* ctx, const nbvh_scene* s, uint32_t precision, const float* d_in,
                                        uint32_t n, float* d_out, float* d_logits, void* stream)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        run_network(ctx, s, precision, d_in, n, d_out, d_logits, (hipStream_t)stream);
    });
}

NBVH_API int32_t nbvh_network_inference_host(nbvh_ctx* ctx, const nbvh_scene* s, uint32_t precision,
                                             const float* in, uint32_t n, float* out, float* logits)
{
    return guarded(ctx, [&] {
        require(ctx && s, "NULL argument");
        if (n == 0) return;
        require(in != nullptr, "NULL input");
        if (!s->has_net) fail(NBVH_ERR_STATE, "network inference on a scene without weights");
        const size_t D = 3 * (size_t)s->cfg.n_points, O = s->cfg.n_outputs;
        DevBuf di, dout, dlog;
        di.upload(ctx, in, (size_t)n * D);
        if (out) dout.alloc(ctx, (size_t)n * O * sizeof(float));
        if (logits) dlog.alloc(ctx, (size_t)n * O * sizeof(float));
        run_network(ctx, s, precision, di.as<float>(), n, dout.as<float>(), dlog.as<float>(), nullptr);
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (out) hip_check(hipMemcpy(out, dout.ptr, (size_t)n * O * sizeof(float), hipMemcpyDeviceToHost), "D2H");
        if (logits) hip_check(hipMemcpy(logits, dlog.ptr, (size_t)n * O * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    });
}
