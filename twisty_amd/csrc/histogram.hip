// histogram.hip -- exit-direction histogram of the path weights (gfx950).
//
// The histogram is build-defined: the reference bins only path-weight
// magnitudes (FullExperimentRunnerOptimalPerturb.cpp:26-37,562-575).  Here the
// weight of every valid path is added to the (cos theta, phi) bin of its last
// free tangent; the result is normalised by the estimator's own sum, so the
// bins add up to 1 and give the angular distribution of the exiting weight.
//
// Path weights are ~1e-40 or far smaller, below FP32 range, so they are binned
// as 2^(log2 w - M) in (0, 1] with M the largest log2 weight of the run (from
// the finalize kernel, read on the device: no host round trip).  A 256x256
// grid is 256 KB of FP32 bins, more than one block's LDS, so pass 1 tiles the
// bins: block (chunk, tile) scans its chunk of paths and adds the ones that
// fall into its tile with LDS float atomics, then writes the tile to
// partial[chunk][bin].  A chunk holds few paths per bin, so these FP32 sums
// are short.  Pass 2 adds the chunks of one bin with compensated (two-sum)
// FP32 and applies the FP64 normaliser 1/S.
#include "device_common.hpp"
#include "forward.hpp"

namespace twisty_hip {

constexpr uint32_t kHistBlock = 1024;
constexpr uint32_t kHistTileBins = 32768;  // 128 KB of LDS per block

__global__ void __launch_bounds__(kHistBlock) hist_tile_kernel(const float2 *pathLog2, const uint32_t *pathBin,
      uint64_t numPaths, uint64_t chunkPaths, const FinalOut *fin, uint32_t numBins, uint32_t tileBins,
      float *partial)
{
    extern __shared__ float tile[];
    const uint32_t t0 = blockIdx.y * tileBins;
    const uint32_t tn = min(tileBins, numBins - t0);
    for (uint32_t i = threadIdx.x; i < tn; i += kHistBlock) tile[i] = 0.0f;
    __syncthreads();
    const float M = (float)fin->maxLog2;
    const uint64_t begin = (uint64_t)blockIdx.x * chunkPaths;
    const uint64_t end = min(numPaths, begin + chunkPaths);
    if (M > -INFINITY && M < INFINITY) {
        for (uint64_t q = begin + threadIdx.x; q < end; q += kHistBlock) {
            const uint32_t b = pathBin[q] - t0;
            if (b < tn) {
                const float2 v = pathLog2[q];
                // NaN (invalid path) fails the comparison; -inf adds zero.
                if (v.x > -INFINITY) atomicAdd(&tile[b], exp2f((v.x - M) + v.y));
            }
        }
    }
    __syncthreads();
    float *out = partial + (uint64_t)blockIdx.x * numBins + t0;
    for (uint32_t i = threadIdx.x; i < tn; i += kHistBlock) out[i] = tile[i];
}

__global__ void __launch_bounds__(256) hist_reduce_kernel(const float *partial, uint32_t numChunks,
      uint32_t numBins, const FinalOut *fin, double *hist)
{
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= numBins) return;
    FF acc { 0.0f, 0.0f };
    for (uint32_t c = 0; c < numChunks; ++c) acc = ff_add_f(acc, partial[(uint64_t)c * numBins + b]);
    const double S = fin->sumScaled;
    hist[b] = (S > 0.0) ? ((double)acc.hi + (double)acc.lo) / S : 0.0;
}

HistShape hist_shape(uint64_t numPaths, uint32_t numBins)
{
    HistShape h {};
    h.numBins = numBins;
    h.tileBins = numBins < kHistTileBins ? numBins : kHistTileBins;
    h.numTiles = (numBins + h.tileBins - 1) / h.tileBins;
    // About 8 paths per thread per chunk, at most 512 chunks.
    uint64_t chunks = (numPaths + 8ull * kHistBlock - 1) / (8ull * kHistBlock);
    if (chunks > 512) chunks = 512;
    if (chunks < 1) chunks = 1;
    h.chunkPaths = (numPaths + chunks - 1) / chunks;
    h.numChunks = (uint32_t)((numPaths + h.chunkPaths - 1) / h.chunkPaths);
    return h;
}

hipError_t launch_histogram(const float2 *pathLog2, const uint32_t *pathBin, uint64_t numPaths,
      const FinalOut *fin, const HistShape &h, float *partial, double *hist, hipStream_t s)
{
    const uint32_t lds = h.tileBins * 4u;
    hipError_t e = hipFuncSetAttribute((const void *)hist_tile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
          (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(hist_tile_kernel, dim3(h.numChunks, h.numTiles), dim3(kHistBlock), lds, s, pathLog2, pathBin,
          numPaths, h.chunkPaths, fin, h.numBins, h.tileBins, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(hist_reduce_kernel, dim3((h.numBins + 255) / 256), dim3(256), 0, s, partial, h.numChunks,
          h.numBins, fin, hist);
    return hipGetLastError();
}

}  // namespace twisty_hip
