// context.h -- the library's context and scene objects and the internal helpers shared by the
// translation units behind the C ABI (api.cpp: scenes and queries; server.cpp: serving loop).
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../nbvh_internal.h"
#include "host_common.h"

#define NBVH_API extern "C" __attribute__((visibility("default")))

struct nbvh_ctx {
    int device = 0;
    int n_cus  = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> used{0};
    mutable std::mutex mu;
    std::string last_error;
    // Per-query scratch slots: queue counters (kCounterWords words) and a part table
    // (kMaxBatchParts entries: device copy + pinned host staging, allocated on first use).
    // Eager queries take slots [0, kEagerSlots) round robin; a slot is handed out again only
    // after the query that last used it has completed (slot_event, waited on by the host).
    // Queries recorded into a HIP graph take a slot of [kEagerSlots, kSlots) each, never reused,
    // so that every replay of the graph finds its counters and part table as captured.
    static constexpr uint32_t kCounterWords  = 32;  // per query: 8 band heads, 16 B apart
    static constexpr uint32_t kEagerSlots    = 2048;
    static constexpr uint32_t kCapturedSlots = 1024;
    static constexpr uint32_t kSlots         = kEagerSlots + kCapturedSlots;
    uint32_t* d_counters = nullptr;  // kSlots x kCounterWords
    nbvh::BatchPart* d_parts = nullptr;
    nbvh::BatchPart* h_parts = nullptr;
    std::atomic<uint32_t> eager_next{0}, captured_next{0};
    hipEvent_t slot_event[kEagerSlots] = {};
    std::mutex slot_mu[kEagerSlots];
};

namespace nbvh {

// Record the failure message on ctx (if any) and on the calling thread.
void set_error(const nbvh_ctx* ctx, const std::string& msg);
// Throws Error(NBVH_ERR_HIP) with the runtime's message when e != hipSuccess.
void hip_check(hipError_t e, const char* what);
// Throws Error(NBVH_ERR_INVALID_ARGUMENT, msg) when cond is false.
void require(bool cond, const char* msg);

// Runs f, turning exceptions into a status and the context's last-error message.
template <typename F>
int32_t guarded(const nbvh_ctx* ctx, F&& f)
{
    try {
        f();
        return NBVH_OK;
    } catch (const Error& e) {
        set_error(ctx, e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_error(ctx, "host allocation failed");
        return NBVH_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error(ctx, e.what());
        return NBVH_ERR_INVALID_ARGUMENT;
    }
}

// Device buffer accounted against the context's HBM budget.
struct DevBuf {
    nbvh_ctx* ctx = nullptr;
    void* ptr     = nullptr;
    size_t bytes  = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release();
    void alloc(nbvh_ctx* c, size_t n);
    template <typename T>
    void upload(nbvh_ctx* c, const T* src, size_t count)
    {
        alloc(c, count * sizeof(T));
        if (count) hip_check(hipMemcpy(ptr, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace nbvh

struct nbvh_scene {
    nbvh_ctx* ctx = nullptr;
    // host copies
    std::vector<float> verts;
    std::vector<uint32_t> tris;
    std::vector<nbvh_node> bvh;
    std::vector<uint32_t> prims;
    std::vector<nbvh_node> nbvh_nodes;
    float bbox_min[3] = {0, 0, 0}, bbox_max[3] = {0, 0, 0};
    bool has_bbox = false, has_net = false;
    uint64_t net_generation = 0;  // bumped whenever the network is replaced (load / init)
    nbvh_network_config cfg{};
    std::vector<float> tables, mlp;
    // device
    nbvh::DevBuf d_verts, d_tris, d_bvh, d_prims, d_nbvh, d_tables32, d_tables16, d_mlp32, d_frag, d_frag8;
    nbvh::NetArgs args{};
    float table_scale = 1.f;
    float fp8_wscale[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    nbvh_launch_shape shape[4] = {};  // per precision (nbvh_autotune / nbvh_scene_set_launch_shape)
};

namespace nbvh {

// Replace the scene's N-BVH (validated; the device copy gets the dilated boxes).
void set_scene_nbvh(nbvh_scene* s, std::vector<nbvh_node> nodes);

// Kernel arguments of a scene from its host state; device copies of its network (all forms).
void refresh_args(nbvh_scene* s);
void upload_network(nbvh_scene* s);

// One query's scratch slot (nbvh_ctx): held from before its first use on the stream until
// finish() has recorded, on that stream, the point after its last use.
class SlotLease {
public:
    SlotLease(nbvh_ctx* ctx, hipStream_t stream);
    SlotLease(const SlotLease&) = delete;
    SlotLease& operator=(const SlotLease&) = delete;
    uint32_t* counters() const;
    BatchPart* host_parts() const;  // requires ensure_part_tables(ctx)
    BatchPart* device_parts() const;
    void finish(hipStream_t stream);  // eager slots: record the completion event
    uint32_t index() const { return k_; }
    bool captured() const { return captured_; }

private:
    nbvh_ctx* ctx_;
    uint32_t k_ = 0;
    bool captured_ = false;
    std::unique_lock<std::mutex> lock_;
};

// One training step (nbvh_trainer_step), optionally writing each sample's loss (the sum of its
// 8 element values) to d_sample_loss; throws Error on failure.
void trainer_step(nbvh_trainer* t, const float* d_inputs, const float* d_targets, uint32_t n, uint32_t flags,
                  float* d_sample_loss, float* loss, hipStream_t stream);
nbvh_scene* trainer_scene(const nbvh_trainer* t);

// Allocates the context's part tables on first use.
void ensure_part_tables(nbvh_ctx* ctx);

// Queue one query launch (counter reset + kernel) over the rays of `batch` on `stream`, with
// the scratch of `lease` (one is taken and finished here when NULL). Validates the parameters
// against the scene; throws Error on failure. scratch: device memory of at least
// query_scratch_bytes(batch.n) for the coherence sort of NBVH_QUERY_SORT_RAYS (NULL: stream-ordered
// allocation).
void run_query(nbvh_ctx* ctx, const nbvh_scene* s, const nbvh_query_params* p, const RayBatch& batch,
               hipStream_t stream, SlotLease* lease = nullptr, void* scratch = nullptr);
size_t query_scratch_bytes(uint32_t n);

// Hybrid closest/any hit of rays [0, min(*live, n)) (all n without live): triangles first, the
// neural query in front of each triangle hit. scratch: at least hybrid_scratch_bytes(n, sort)
// (NULL: stream-ordered allocation).
void run_hybrid(nbvh_ctx* ctx, const nbvh_scene* neural, const nbvh_scene* triangles, const nbvh_query_params* p,
                const nbvh_ray* d_rays, uint32_t n, nbvh_hit* d_hits, hipStream_t st, const uint32_t* live = nullptr,
                void* scratch = nullptr);
size_t hybrid_scratch_bytes(uint32_t n, bool sort);

}  // namespace nbvh
