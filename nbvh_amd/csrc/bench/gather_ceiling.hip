// gather_ceiling.hip -- measurement helper of the benchmark (not part of the query library):
// the chip-wide rate at which random 128-byte lines can be gathered from a table of a given size.
//
// Every lane of a wave loads one 4-byte word; `share` consecutive lanes read the same 128-byte
// line, so a load instruction touches 64 / share distinct lines, chosen by a hash of (wave,
// iteration, lane / share) over the first power-of-two number of lines of the table. Each lane
// issues its loads in groups of 8 before using them. The rate in lines per second is the ceiling
// the hash-grid gathers of the neural query run against: about 254 G lines/s while the table fits
// one XCD's L2 and about 60 G lines/s (HBM) beyond.
//
// nbvh_bench_atomic_adds measures the rate of random FP32 atomic adds, the ceiling of the training
// step's table-gradient scatter.
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

// Pairs of FP32 atomic adds (no return value) of 1.0 at random entries of a table of F = 2 floats
// per entry, the access pattern of the hash-grid gradient scatter: each lane adds to both floats of
// `per_lane` entries chosen by a hash of (thread, iteration).
__global__ __launch_bounds__(256) void atomic_pairs_kernel(float* __restrict__ tab, uint32_t entries,
                                                           uint32_t per_lane)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < per_lane; k++) {
        const uint32_t e = mix32(tid * 1021u + k * 7919u + 17u) % entries;
        atomicAdd(tab + 2 * (size_t)e, 1.f);
        atomicAdd(tab + 2 * (size_t)e + 1, 1.f);
    }
}

__global__ __launch_bounds__(256) void sum_kernel(const float* __restrict__ tab, uint64_t n, double* __restrict__ out)
{
    double s = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += tab[i];
    atomicAdd(out, s);
}

}  // namespace

// Returns 0 and the mean time of one launch (ms), the rate of lane atomic adds per second and
// whether the table's sum equals the number of adds (*exact = 1), or a HIP error code.
// entries in [1, 2^28]; waves_per_cu a multiple of 4; per_lane >= 1.
extern "C" __attribute__((visibility("default"))) int nbvh_bench_atomic_adds(uint32_t entries, uint32_t waves_per_cu,
                                                                              uint32_t per_lane, float* ms,
                                                                              double* adds_per_s, int32_t* exact)
{
    if (entries == 0 || entries > (1u << 28) || waves_per_cu == 0 || waves_per_cu % 4 || per_lane == 0)
        return (int)hipErrorInvalidValue;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    const uint32_t blocks = (uint32_t)cus * waves_per_cu / 4;
    const uint64_t floats = 2ull * entries;
    float* tab = nullptr;
    double* sum = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    e = hipMalloc((void**)&tab, floats * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&sum, sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const int reps = 5;
    float total = 0.f;
    double s = 0.0;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipMemset(tab, 0, floats * 4);
        if (e == hipSuccess) e = hipMemset(sum, 0, sizeof(double));
        if (e == hipSuccess) e = hipEventRecord(a, 0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(atomic_pairs_kernel, dim3(blocks), dim3(256), 0, 0, tab, entries, per_lane);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(b, 0);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(sum_kernel, dim3(1024), dim3(256), 0, 0, (const float*)tab, floats, sum);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
        if (e == hipSuccess) e = hipMemcpy(&s, sum, sizeof(double), hipMemcpyDeviceToHost);
        total += t;
    }
    if (e == hipSuccess) {
        const double adds = (double)blocks * 256 * per_lane * 2;
        *ms = total / reps;
        *adds_per_s = adds / (*ms * 1e-3);
        *exact = s == adds ? 1 : 0;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (sum) (void)hipFree(sum);
    if (tab) (void)hipFree(tab);
    return (int)e;
}

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
