// On-box HBM reference for the benchmark: a streaming copy kernel (16-byte non-temporal loads
// and stores, grid-stride) timed with HIP events. A measurement helper of bench.py only; not
// part of the solver library or its C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef float V4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void copyKernel(const V4* __restrict__ src, V4* __restrict__ dst, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace

// Copy bandwidth in GB/s (read + write bytes per second) of `bytes` per copy, the mean over
// `reps` copies after one warm-up; negative on a HIP error.
extern "C" double dgprobe_copy_gbps(int32_t device, int64_t bytes, int32_t reps) {
    if (bytes < 16 || reps < 1) return -1.0;
    if (hipSetDevice(device) != hipSuccess) return -1.0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return -1.0;
    const int64_t n = bytes / 16;
    V4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return -1.0;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return -1.0;
    }
    double out = -1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const dim3 grid(static_cast<unsigned>(cus * 4)), block(256);  // 4 blocks per CU: the fastest of 1..32 measured
    if (hipMemset(a, 0, n * 16) == hipSuccess && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        hipLaunchKernelGGL(copyKernel, grid, block, 0, nullptr, a, b, n);
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copyKernel, grid, block, 0, nullptr, a, b, n);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess &&
            hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f)
            out = 2.0 * static_cast<double>(n * 16) * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return out;
}
