// valu_rate.hip -- issue cost of the VALU instructions the path kernels lean
// on, measured on the device: every thread runs 8 independent chains of one
// operation (no dependency stalls at 32 waves per CU), so the time per
// wave-instruction per SIMD is the instruction's issue cost.  Used to read the
// kernels' SQ_INSTS_VALU counts against what the SIMDs can issue.
// Usage: twisty_valu_rate [reps]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kIter = 4096;   // loop iterations per thread
constexpr int kChains = 8;    // independent chains per thread
constexpr int kBlocks = 256 * 32;

template<int OP>
__global__ void __launch_bounds__(256) rate_kernel(float *out, uint32_t seed)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    float f[kChains];
    uint32_t u[kChains], h[kChains];
    for (int c = 0; c < kChains; ++c) {
        f[c] = 1.0f + (float)((t + c) & 7) * 1e-3f;
        u[c] = t * 2654435761u + c + seed;
        h[c] = u[c] ^ 0x9E3779B9u;
    }
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if constexpr (OP == 0) f[c] = __builtin_fmaf(f[c], 0.999f, 1e-3f);            // v_fma_f32
            if constexpr (OP == 1) f[c] = __builtin_amdgcn_rsqf(f[c]);                    // v_rsq_f32
            if constexpr (OP == 2) f[c] = __builtin_amdgcn_sqrtf(f[c]);                   // v_sqrt_f32
            if constexpr (OP == 3) f[c] = __builtin_amdgcn_logf(f[c]) + 1.5f;             // v_log_f32 + v_add_f32
            if constexpr (OP == 4) {                                                      // v_mad_u64_u32 + xor
                const uint64_t p = (uint64_t)u[c] * 0xD2511F53u;
                u[c] = (uint32_t)(p >> 32) ^ h[c];
                h[c] = (uint32_t)p;
            }
        }
    }
    float s = 0.0f;
    for (int c = 0; c < kChains; ++c) s += f[c] + (float)(u[c] ^ h[c]);
    out[t] = s;  // t < kBlocks * 256, the size of out
}

template<int OP>
float time_op(float *out, int reps)
{
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int r = 0; r < reps + 1; ++r) {  // the first launch is a warm-up
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(rate_kernel<OP>, dim3(kBlocks), dim3(256), 0, 0, out, (uint32_t)r);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) std::exit(1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

int main(int argc, char **argv)
{
    const int reps = argc > 1 ? std::atoi(argv[1]) : 3;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * kBlocks * 256) != hipSuccess) return 1;
    const double simds = 4.0 * prop.multiProcessorCount;
    const double perSimd = (double)kBlocks * 4 * kIter * kChains / simds;  // wave-instructions of the op per SIMD
    const char *names[] = { "v_fma_f32", "v_rsq_f32", "v_sqrt_f32", "v_log_f32 + v_add_f32",
                            "v_mad_u64_u32 + v_xor_b32" };
    const float ms[] = { time_op<0>(out, reps), time_op<1>(out, reps), time_op<2>(out, reps), time_op<3>(out, reps),
                         time_op<4>(out, reps) };
    std::printf("%d CUs; %d blocks of 256 threads, %d iterations of %d chains each; best of %d launches\n",
                prop.multiProcessorCount, kBlocks, kIter, kChains, reps);
    for (int k = 0; k < 5; ++k)
        std::printf("%-28s %8.3f ms  %.3f ns per wave-instruction per SIMD  %.2f x v_fma_f32\n", names[k], ms[k],
                    ms[k] * 1e6 / perSimd, ms[k] / ms[0]);
    (void)hipFree(out);
    return 0;
}
