// noalloc.hip -- the statically allocated scratch of nbvh_query_noalloc.
//
// A query's scratch is its queue counters (8 band heads, 16 B apart) and, for the coherence-
// sorted form, the sort keys, ray ids, the radix sort's temporary storage and the claim order.
// nbvh_query takes them from the context (counters) and the stream-ordered allocator (sort).
// nbvh_query_noalloc takes them from one file-scope device array instead, so it calls no
// allocator: usable where hipMalloc is not (graph capture, allocation-free serving loops).
//
// The array exists once per GPU the library runs on and lives until the process ends. Host code
// must never pass `scratch` itself to a kernel: in host code the name is a host-side stand-in,
// not a device address. Its device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept in g_addr.
#include <hip/hip_runtime.h>

#include <mutex>

#include "launch.h"

namespace nbvh {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_addr[device] != nullptr) return hipSuccess;
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_addr[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_addr[device];
}

}  // namespace nbvh
