// copy_calib.hip -- plain device-to-device copy of a known size, used to scale
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters into bytes for the roofline's
// traffic figure.  Usage: twisty_copy_calib [MiB] [reps]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char **argv)
{
    const size_t mib = argc > 1 ? (size_t)std::atol(argv[1]) : 256;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 3;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        std::printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms, 2.0 * bytes / (ms * 1e6));
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
