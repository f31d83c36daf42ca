// server.cpp -- continuous batching of ray-query requests over device-resident scenes.
//
// The reference renders one frame at a time from a single host thread
// (NeuralPathModule::neural_bvh_traversal, src/cuda/neural/path/path_module.cu:54-137): one
// batch, one stream, blocking syncs. Here any number of client threads submit requests (a ray
// array, a scene, query parameters) and wait on tickets. One scheduler thread keeps up to
// n_streams launches in flight on their own streams, so transfers and kernels of different
// launches overlap. Whenever a stream frees up it takes the oldest queued request and every
// later queued request with the same scene and parameters (closest- and any-hit requests of one
// family mix: the any-hit flag is per part; up to kMaxBatchParts requests and max_batch_rays
// rays) and runs them as ONE persistent launch over a part table: requests that
// arrive while others are on the GPU join the next launch. Requests whose rays are in host
// memory are staged through a fixed device arena; a request waits until its 80 B per ray fit
// (admission by HBM budget), and requests larger than the arena are refused at submit.
//
// Every ray's result depends on that ray alone (the kernels evaluate each segment's network
// column independently, in a fixed order), so a request's hits are identical whether it runs
// alone or inside a mixed launch.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "context.h"

namespace nbvh {
namespace {

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr uint64_t kStageBytesPerRay = sizeof(nbvh_ray) + sizeof(nbvh_hit);
constexpr uint64_t kStageAlign       = 256;

// First-fit allocator over the staging arena (offsets in bytes).
class Arena {
public:
    void reset(uint64_t size)
    {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
    }
    uint64_t size() const { return size_; }
    bool alloc(uint64_t n, uint64_t* off)
    {
        n = (n + kStageAlign - 1) / kStageAlign * kStageAlign;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n)
    {
        n = (n + kStageAlign - 1) / kStageAlign * kStageAlign;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }

private:
    std::map<uint64_t, uint64_t> free_;  // offset -> size
    uint64_t size_ = 0;
};

struct Request {
    uint64_t ticket = 0;
    const nbvh_scene* scene = nullptr;
    nbvh_query_params params{};
    const nbvh_ray* rays = nullptr;  // caller memory (host or device)
    nbvh_hit* hits       = nullptr;
    uint32_t n           = 0;
    bool device          = false;
    // staging (host requests)
    bool staged = false;
    uint64_t stage_off = 0;
    // outcome
    bool complete = false;
    nbvh_request_info info{};
    std::string error;
};

struct Launch;

// One stream with its part table (device + pinned host copy); one launch in flight at a time.
struct StreamSlot {
    hipStream_t stream   = nullptr;
    BatchPart* d_parts   = nullptr;
    BatchPart* h_parts   = nullptr;  // pinned
    bool busy            = false;
};

struct Launch {
    struct nbvh_server* server = nullptr;
    uint32_t slot = 0;
    std::vector<std::shared_ptr<Request>> reqs;
    bool done = false;
    uint64_t done_ns = 0;
};

bool is_triangle(uint32_t kind) { return kind == NBVH_QUERY_TRIANGLE_CLOSEST || kind == NBVH_QUERY_TRIANGLE_ANY; }
bool is_any_hit(uint32_t kind) { return kind == NBVH_QUERY_NEURAL_ANY || kind == NBVH_QUERY_TRIANGLE_ANY; }

// Requests that can share a launch: same scene and kind family (closest and any hit mix, per
// part), same precision, lod, step limit and threshold.
bool same_key(const Request& a, const Request& b)
{
    return a.scene == b.scene && is_triangle(a.params.kind) == is_triangle(b.params.kind) &&
           a.params.precision == b.params.precision && a.params.lod == b.params.lod &&
           a.params.max_traversal_steps == b.params.max_traversal_steps &&
           a.params.visibility_threshold == b.params.visibility_threshold &&
           a.params.segment_seed == b.params.segment_seed;
}

}  // namespace
}  // namespace nbvh

using namespace nbvh;

struct nbvh_server {
    nbvh_ctx* ctx = nullptr;
    nbvh_server_config cfg{};
    std::vector<StreamSlot> slots;
    DevBuf d_stage;
    Arena arena;

    std::mutex mu;
    std::condition_variable cv_sched;  // scheduler: new work, finished launch, stop
    std::condition_variable cv_done;   // waiters: a request completed
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;  // submitted, not yet waited for
    std::vector<std::unique_ptr<Launch>> inflight;
    uint64_t next_ticket = 1;
    bool stop = false;
    nbvh_server_stats stats{};
    std::thread thread;

    void run();
    void reap_locked();
    bool form_batch_locked(Launch& L);
    void launch(Launch& L);
    void finish_locked(Launch& L, int32_t status, const std::string& msg);
};

namespace {

// Stream-ordered completion callback (runs on a runtime thread; no HIP calls here).
void on_launch_done(void* p)
{
    Launch* L = static_cast<Launch*>(p);
    nbvh_server* s = L->server;
    std::lock_guard<std::mutex> g(s->mu);
    L->done    = true;
    L->done_ns = now_ns();
    s->cv_sched.notify_all();
}

}  // namespace

void nbvh_server::finish_locked(Launch& L, int32_t status, const std::string& msg)
{
    for (auto& r : L.reqs) {
        if (r->staged) {
            arena.release(r->stage_off, (uint64_t)r->n * kStageBytesPerRay);
            r->staged = false;
        }
        r->info.status      = status;
        r->info.complete_ns = L.done_ns ? L.done_ns : now_ns();
        r->complete         = true;
        stats.n_completed++;
        if (status != NBVH_OK) {
            stats.n_failed++;
            r->error = msg;
        }
    }
    slots[L.slot].busy = false;
    cv_done.notify_all();
}

void nbvh_server::reap_locked()
{
    for (size_t i = 0; i < inflight.size();) {
        if (inflight[i]->done) {
            finish_locked(*inflight[i], NBVH_OK, "");
            inflight.erase(inflight.begin() + (long)i);
        } else {
            i++;
        }
    }
}

// Oldest queued request, then every later one with the same key, within the launch limits and
// the staging arena. Requests are admitted in arrival order per key.
bool nbvh_server::form_batch_locked(Launch& L)
{
    if (pending.empty()) return false;
    const Request& head = *pending.front();
    uint64_t rays = 0;
    for (auto it = pending.begin(); it != pending.end() && L.reqs.size() < kMaxBatchParts;) {
        Request& r = **it;
        if (!same_key(r, head)) {
            ++it;
            continue;
        }
        if (!L.reqs.empty() && rays + r.n > cfg.max_batch_rays) break;
        if (!r.device) {
            uint64_t off = 0;
            if (!arena.alloc((uint64_t)r.n * kStageBytesPerRay, &off)) break;
            r.staged    = true;
            r.stage_off = off;
        }
        rays += r.n;
        L.reqs.push_back(*it);
        it = pending.erase(it);
    }
    return !L.reqs.empty();
}

// Queue transfers, the part table, the query and the completion callback on the slot's stream.
void nbvh_server::launch(Launch& L)
{
    StreamSlot& sl = slots[L.slot];
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    uint32_t begin = 0;
    char* stage = static_cast<char*>(d_stage.ptr);
    for (size_t k = 0; k < L.reqs.size(); k++) {
        Request& r = *L.reqs[k];
        BatchPart& part = sl.h_parts[k];
        if (r.device) {
            part.rays = r.rays;
            part.hits = r.hits;
        } else {
            nbvh_ray* dr = reinterpret_cast<nbvh_ray*>(stage + r.stage_off);
            nbvh_hit* dh = reinterpret_cast<nbvh_hit*>(stage + r.stage_off + (uint64_t)r.n * sizeof(nbvh_ray));
            hip_check(hipMemcpyAsync(dr, r.rays, (size_t)r.n * sizeof(nbvh_ray), hipMemcpyHostToDevice, sl.stream),
                      "request rays H2D");
            part.rays = dr;
            part.hits = dh;
        }
        part.begin    = begin;
        part.n        = r.n;
        part.any_hit  = is_any_hit(r.params.kind) ? 1u : 0u;
        part.reserved = 0u;
        begin += r.n;
        r.info.launch_ns   = now_ns();
        r.info.n_launches  = 1;
        r.info.batch_parts = (uint32_t)L.reqs.size();
    }
    const uint32_t n_parts = (uint32_t)L.reqs.size();
    hip_check(hipMemcpyAsync(sl.d_parts, sl.h_parts, n_parts * sizeof(BatchPart), hipMemcpyHostToDevice, sl.stream),
              "part table H2D");
    const RayBatch batch{nullptr, nullptr, sl.d_parts, n_parts, begin};
    run_query(ctx, L.reqs[0]->scene, &L.reqs[0]->params, batch, sl.stream);
    for (size_t k = 0; k < L.reqs.size(); k++) {
        Request& r = *L.reqs[k];
        if (r.device) continue;
        hip_check(hipMemcpyAsync(r.hits, sl.h_parts[k].hits, (size_t)r.n * sizeof(nbvh_hit), hipMemcpyDeviceToHost,
                                 sl.stream),
                  "request hits D2H");
    }
    hip_check(hipLaunchHostFunc(sl.stream, on_launch_done, &L), "completion callback");
}

void nbvh_server::run()
{
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        reap_locked();
        if (stop && pending.empty() && inflight.empty()) break;
        uint32_t free_slot = UINT32_MAX;
        for (uint32_t i = 0; i < slots.size(); i++)
            if (!slots[i].busy) {
                free_slot = i;
                break;
            }
        if (free_slot != UINT32_MAX && !pending.empty()) {
            std::unique_ptr<Launch> L(new Launch);
            L->server = this;
            L->slot   = free_slot;
            if (form_batch_locked(*L)) {
                slots[free_slot].busy = true;
                Launch* raw = L.get();
                inflight.push_back(std::move(L));
                stats.n_launches++;
                stats.n_batched_requests += raw->reqs.size();
                lk.unlock();
                std::string err;
                int32_t status = NBVH_OK;
                try {
                    launch(*raw);
                } catch (const Error& e) {
                    status = e.status;
                    err    = e.what();
                } catch (const std::exception& e) {
                    status = NBVH_ERR_INVALID_ARGUMENT;
                    err    = e.what();
                }
                lk.lock();
                if (status != NBVH_OK) {
                    // Nothing of this launch is pending on the stream except possibly copies
                    // queued before the failure; wait for them before the staging is reused.
                    lk.unlock();
                    (void)hipStreamSynchronize(slots[raw->slot].stream);
                    lk.lock();
                    raw->done_ns = now_ns();
                    finish_locked(*raw, status, err);
                    inflight.erase(std::find_if(inflight.begin(), inflight.end(),
                                                [&](const std::unique_ptr<Launch>& p) { return p.get() == raw; }));
                }
                continue;
            }
            // Nothing admissible (the head request waits for staging space): wait for a launch.
        }
        cv_sched.wait(lk);
    }
}

// =============================================================================================
// extern "C"
// =============================================================================================

NBVH_API int32_t nbvh_server_create(nbvh_ctx* ctx, const nbvh_server_config* cfg, nbvh_server** out)
{
    return guarded(ctx, [&] {
        require(ctx && out, "NULL argument");
        *out = nullptr;
        std::unique_ptr<nbvh_server> s(new nbvh_server);
        s->ctx = ctx;
        if (cfg) s->cfg = *cfg;
        if (s->cfg.n_streams == 0) s->cfg.n_streams = 4;
        if (s->cfg.max_batch_rays == 0) s->cfg.max_batch_rays = 1u << 21;
        if (s->cfg.staging_bytes == 0) s->cfg.staging_bytes = 256ull << 20;
        require(s->cfg.n_streams <= 16, "n_streams must be at most 16");
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        s->d_stage.alloc(ctx, s->cfg.staging_bytes);
        s->arena.reset(s->cfg.staging_bytes);
        s->slots.resize(s->cfg.n_streams);
        struct Cleanup {
            nbvh_server* s;
            bool armed = true;
            ~Cleanup()
            {
                if (!armed) return;
                for (auto& sl : s->slots) {
                    if (sl.stream) (void)hipStreamDestroy(sl.stream);
                    if (sl.d_parts) (void)hipFree(sl.d_parts);
                    if (sl.h_parts) (void)hipHostFree(sl.h_parts);
                }
            }
        } cleanup{s.get()};
        for (auto& sl : s->slots) {
            hip_check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipMalloc((void**)&sl.d_parts, kMaxBatchParts * sizeof(BatchPart)), "hipMalloc part table");
            hip_check(hipHostMalloc((void**)&sl.h_parts, kMaxBatchParts * sizeof(BatchPart), hipHostMallocDefault),
                      "hipHostMalloc part table");
        }
        cleanup.armed = false;
        nbvh_server* raw = s.get();
        raw->thread = std::thread([raw] { raw->run(); });
        *out = s.release();
    });
}

NBVH_API int32_t nbvh_server_destroy(nbvh_server* s)
{
    if (!s) return NBVH_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_sched.notify_all();
    }
    if (s->thread.joinable()) s->thread.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto& sl : s->slots) {
        (void)hipStreamSynchronize(sl.stream);
        (void)hipStreamDestroy(sl.stream);
        (void)hipFree(sl.d_parts);
        (void)hipHostFree(sl.h_parts);
    }
    delete s;
    return NBVH_OK;
}

NBVH_API int32_t nbvh_request_submit(nbvh_server* s, const nbvh_scene* scene, const nbvh_query_params* params,
                                     const nbvh_ray* rays, uint32_t n_rays, nbvh_hit* hits, uint32_t flags,
                                     uint64_t* ticket)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && scene && params && ticket, "NULL argument");
        require(scene->ctx == s->ctx, "scene belongs to another context");
        require(n_rays > 0, "empty request");
        require(rays && hits, "NULL ray or hit buffer");
        require((flags & ~(uint32_t)NBVH_REQUEST_DEVICE_BUFFERS) == 0u, "unknown request flags");
        require(params->kind <= NBVH_QUERY_TRIANGLE_ANY, "unknown query kind");
        require(params->precision <= NBVH_PREC_FP8, "unknown precision");
        const bool device = (flags & NBVH_REQUEST_DEVICE_BUFFERS) != 0u;
        if (!device && (uint64_t)n_rays * kStageBytesPerRay > s->cfg.staging_bytes)
            fail(NBVH_ERR_CAPACITY, "request of " + std::to_string(n_rays) +
                                        " host rays exceeds the server's staging arena");
        auto r = std::make_shared<Request>();
        r->scene  = scene;
        r->params = *params;
        r->rays   = rays;
        r->hits   = hits;
        r->n      = n_rays;
        r->device = device;
        r->info.submit_ns = now_ns();
        std::lock_guard<std::mutex> g(s->mu);
        require(!s->stop, "server is shutting down");
        r->ticket = s->next_ticket++;
        s->live[r->ticket] = r;
        s->pending.push_back(r);
        s->stats.n_submitted++;
        s->stats.n_rays += n_rays;
        *ticket = r->ticket;
        s->cv_sched.notify_all();
    });
}

NBVH_API int32_t nbvh_request_wait(nbvh_server* s, uint64_t ticket, int64_t timeout_us, nbvh_request_info* info)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s != nullptr, "NULL server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->live.find(ticket);
        if (it == s->live.end()) fail(NBVH_ERR_INVALID_ARGUMENT, "unknown or already collected ticket");
        std::shared_ptr<Request> r = it->second;
        auto ready = [&] { return r->complete; };
        if (timeout_us < 0)
            s->cv_done.wait(lk, ready);
        else if (!s->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), ready))
            fail(NBVH_ERR_TIMEOUT, "request not complete");
        s->live.erase(ticket);
        if (info) *info = r->info;
        if (r->info.status != NBVH_OK) fail(r->info.status, "request failed: " + r->error);
    });
}

NBVH_API int32_t nbvh_server_get_stats(nbvh_server* s, nbvh_server_stats* out)
{
    return guarded(s ? s->ctx : nullptr, [&] {
        require(s && out, "NULL argument");
        std::lock_guard<std::mutex> g(s->mu);
        *out = s->stats;
        out->n_pending  = (uint32_t)s->pending.size();
        out->n_inflight = (uint32_t)s->inflight.size();
    });
}
