// gather_ceiling.hip -- measurement helper of the benchmark (not part of the query library):
// the chip-wide rate at which random 128-byte lines can be gathered from a table of a given size.
//
// Every lane of a wave loads one 4-byte word; `share` consecutive lanes read the same 128-byte
// line, so a load instruction touches 64 / share distinct lines, chosen by a hash of (wave,
// iteration, lane / share) over the first power-of-two number of lines of the table. Each lane
// issues its loads in groups of 8 before using them. The rate in lines per second is the ceiling
// the hash-grid gathers of the neural query run against: about 254 G lines/s while the table fits
// one XCD's L2 and about 60 G lines/s (HBM) beyond.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(256) void gather_lines_kernel(const uint32_t* __restrict__ tab, uint32_t line_mask,
                                                           uint32_t share, uint32_t iters, uint32_t* __restrict__ out)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; it += 8) {
        uint32_t v[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) {
            const uint32_t line = mix32(wave * 7919u + (it + u) * 104729u + lane / share) & line_mask;
            v[u] = tab[line * 32u + lane % share];
        }
#pragma unroll
        for (uint32_t u = 0; u < 8; u++) acc ^= v[u];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

// Returns 0 and the mean time of one launch (ms) and the line rate (lines/s), or a HIP error code.
// table_bytes >= 128; share in {1, 2, 4, ..., 32}; iters a multiple of 8.
extern "C" __attribute__((visibility("default"))) int nbvh_bench_gather_lines(uint64_t table_bytes, uint32_t share,
                                                                               uint32_t waves_per_cu, uint32_t iters,
                                                                               float* ms, double* lines_per_s)
{
    if (table_bytes < 128 || share == 0 || share > 32 || (share & (share - 1)) || iters == 0 || iters % 8 ||
        waves_per_cu == 0 || waves_per_cu % 4)
        return (int)hipErrorInvalidValue;
    uint64_t lines = table_bytes / 128, p2 = 1;
    while (p2 * 2 <= lines) p2 *= 2;
    if (p2 > (1ull << 31)) return (int)hipErrorInvalidValue;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    const uint32_t blocks = (uint32_t)cus * waves_per_cu / 4;
    uint32_t *tab = nullptr, *out = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    e = hipMalloc((void**)&tab, p2 * 128);
    if (e == hipSuccess) e = hipMemset(tab, 0x5a, p2 * 128);
    if (e == hipSuccess) e = hipMalloc((void**)&out, (size_t)blocks * 256 * 4);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const int reps = 5;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather_lines_kernel, dim3(blocks), dim3(256), 0, 0, tab, (uint32_t)(p2 - 1), share, iters,
                           out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(a, 0);
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(gather_lines_kernel, dim3(blocks), dim3(256), 0, 0, tab, (uint32_t)(p2 - 1), share, iters,
                           out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (e == hipSuccess) {
        *ms = t / reps;
        *lines_per_s = (double)blocks * 4 * iters * (64.0 / share) / (*ms * 1e-3);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (out) (void)hipFree(out);
    if (tab) (void)hipFree(tab);
    return (int)e;
}
