// calib_copy.hip -- a plain device-to-device copy of known size (1 GiB each
// way), profiled under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to find the
// factor between those counters and the bytes a kernel really moves.  A
// measurement aid for the roofline figures; not part of libvfb.so.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void k_copy(const unsigned long long* __restrict__ a, unsigned long long* __restrict__ b, size_t n) {
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += s) b[i] = a[i];
}

// the same copy with 16-byte accesses, four loads in flight per thread before
// the stores: the HBM rate a streaming kernel can reach on this device
// (n16 = 16-byte elements, a multiple of 4 * gridDim.x * blockDim.x)
__global__ __launch_bounds__(256) void k_copy_wide(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n16) {
    const size_t s = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += 4 * s) {
        uint4 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = a[i + j * s];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[i + j * s] = v[j];
    }
}

int main() {
    const size_t n = (size_t)1 << 27;   // 2^27 words = 1 GiB
    unsigned long long *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * 8) != hipSuccess || hipMalloc((void**)&b, n * 8) != hipSuccess) {
        std::fprintf(stderr, "calib_copy: hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, n * 8) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    float best = 1e30f;
    for (int r = 0; r < 4; ++r) {
        if (hipEventRecord(e0) != hipSuccess) return 1;
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        if (r && ms < best) best = ms;
    }
    std::printf("k_copy: %zu bytes read + %zu written per launch, best %.3f ms = %.1f GB/s\n", n * 8, n * 8, best,
                2.0 * n * 8 / best / 1e6);
    // grid: 256 CUs x 8 workgroups of 256 threads; n16 = 2^26 is a multiple of 4 * 2^19
    const unsigned grid = 2048;
    const size_t n16 = n / 2;
    if ((n16 % (4ull * grid * 256)) != 0) return 1;
    float bestw = 1e30f;
    for (int r = 0; r < 6; ++r) {
        if (hipEventRecord(e0) != hipSuccess) return 1;
        hipLaunchKernelGGL(k_copy_wide, dim3(grid), dim3(256), 0, 0, (const uint4*)a, (uint4*)b, n16);
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        if (r && ms < bestw) bestw = ms;
    }
    std::printf("k_copy_wide: %zu bytes read + %zu written per launch, best %.3f ms = %.1f GB/s\n", n * 8, n * 8,
                bestw, 2.0 * n * 8 / bestw / 1e6);
    return hipFree(a) == hipSuccess && hipFree(b) == hipSuccess ? 0 : 1;
}
