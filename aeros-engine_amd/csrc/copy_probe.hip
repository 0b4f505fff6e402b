// A plain device copy of known size, used only to scale the rocprofv3
// FETCH_SIZE / WRITE_SIZE counters into bytes: the step kernel's counter
// readings are multiplied by (bytes copied) / (counter reading of this copy).
// Not part of the solver's C ABI.
#include <hip/hip_runtime.h>

#include <stddef.h>

namespace {
__global__ void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}
}  // namespace

// Copies `bytes` (a multiple of 16) between two fresh device buffers `reps`
// times; returns 0 on success.
extern "C" int aeros_probe_copy(size_t bytes, int reps) {
    if (bytes == 0 || bytes % 16 != 0 || reps <= 0) return 1;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return 2;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return 2; }
    int rc = 0;
    const size_t n4 = bytes / 16;
    const unsigned grid = (unsigned)((n4 + 255) / 256);
    if (hipMemset(a, 0, bytes) != hipSuccess) rc = 3;
    for (int r = 0; r < reps && !rc; ++r) {
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, 0, (const float4*)a, (float4*)b, n4);
        if (hipGetLastError() != hipSuccess) rc = 4;
    }
    if (hipDeviceSynchronize() != hipSuccess) rc = 5;
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
