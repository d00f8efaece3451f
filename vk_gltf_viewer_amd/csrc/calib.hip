// Calibration of the visibility-buffer write: the device's throughput for
// 64-bit atomicMin without a returned value (the raster's fragment write),
// on a 3840 x 2160 key buffer, for two access patterns:
//   coalesced  the 64 lanes of a wave update 64 consecutive keys (512 bytes)
//   scattered  every lane updates a pseudo-random key
// It is the ceiling the raster's atomics per second are reported against.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.hpp"

namespace vkgv {
namespace {

constexpr uint32_t kCalibBlock = 256;
constexpr uint64_t kCalibKeys = 3840ull * 2160ull;

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(kCalibBlock) void atomic_rate_kernel(unsigned long long* keys, uint32_t iters,
                                                                  uint32_t scattered)
{
    const uint32_t tid = blockIdx.x * kCalibBlock + threadIdx.x;
    const uint32_t nthreads = gridDim.x * kCalibBlock;
    for (uint32_t k = 0; k < iters; ++k) {
        const uint32_t h = mix32(tid * 0x9E3779B9u + k);
        const uint64_t i = scattered ? (uint64_t)h % kCalibKeys : ((uint64_t)k * nthreads + tid) % kCalibKeys;
        (void)__hip_atomic_fetch_min(keys + i, ((unsigned long long)h << 32) | tid, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    }
}

} // namespace

// Atomics per second of the best of three timed launches (the stream is
// synchronized).  `keys` holds kCalibKeys keys.
int measure_atomic_rate(unsigned long long* keys, uint32_t cus, uint32_t scattered, hipStream_t st, double* rate)
{
    const uint32_t blocks = cus * 8u;
    const uint32_t iters = 256u;
    const double atomics = (double)blocks * kCalibBlock * iters;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return -1; }
    float best = 0.0f;
    bool ok = true;
    atomic_rate_kernel<<<blocks, kCalibBlock, 0, st>>>(keys, iters, scattered);   // warm-up
    for (int r = 0; r < 3 && ok; ++r) {
        ok = hipEventRecord(e0, st) == hipSuccess;
        atomic_rate_kernel<<<blocks, kCalibBlock, 0, st>>>(keys, iters, scattered);
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float ms = 0.0f;
        ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && (best == 0.0f || ms < best)) best = ms;
    }
    ok = ok && hipGetLastError() == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok || best <= 0.0f) return -1;
    *rate = atomics / (best * 1e-3);
    return 0;
}

uint64_t calib_key_count() { return kCalibKeys; }

} // namespace vkgv
