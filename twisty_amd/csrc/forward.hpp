// forward.hpp -- kernel parameter blocks and launchers shared by the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace twisty_hip {

struct FwdParams {
    const float *initSeg;  // N segment vectors p_{i+1} - p_i, x,y,z floats
    const float *table;    // tableSize floats (device)
    BlockPartial *partials;
    float *perPathLog10;   // nullable, numChains*chainLength floats
    float *perJointCurv;   // nullable, numChains*chainLength*(N-1) floats
    uint64_t seed;
    uint64_t pathOffset;
    uint64_t numChains;    // ceil(numPaths / chainLength)
    uint64_t numPaths;
    uint32_t chainLength;
    uint32_t burnIn;       // paths perturbed but not weighted at the start of every chain
    uint32_t N;
    uint32_t K;
    uint32_t maxDiff;
    uint32_t stream;
    uint32_t S;            // segment slots per lane = ceil(N / G)
    uint32_t scratchRow;   // index of the lane's zero row: max(S, window - 1)
    int32_t tableSize;
    uint32_t tableStride;  // floats reserved for the table in LDS (multiple of 64)
    float minK, maxK, step, invDs;
    float log10of2;
    float absorbLog10PerPath;
    // Optional per-path outputs of the forms built on the forward pass.
    float2 *pathLog2;      // nullable: (hi, lo) of log2 w without absorption; hi NaN = invalid
    uint32_t *pathBin;     // nullable: exit-direction bin of the last free tangent t_{N-2}
    float *finalSeg;       // nullable: final segment i, component c of path local at (3i + c)*numPaths + local
    float exitE[3], exitF1[3], exitF2[3];  // end frame of the exit-direction bins
    uint32_t nTheta, nPhi;
    // Weight-magnitude histogram (nullable): path counts over weightBins linear
    // bins of 10^(log10 w + normaliser) on [weightMin, weightMax].
    unsigned long long *weightHist;
    uint32_t weightBins;
    uint32_t histStride;   // LDS words reserved for the block's bins (multiple of 64, 0 when off)
    double weightMin, weightMax;
    double weightLog10Offset;  // absorption per path + path normaliser, log10
};

struct FinalOut {
    double log10SumW;
    double log10SumW2;
    double maxLog10W;
    unsigned long long nValid, nInvalid, nClamped;
    double maxLog2;     // M: max per-path log2 weight (no absorption), -inf if none
    double sumScaled;   // S = sum over valid paths of 2^(log2 w - M)
};

// Where a forward instance keeps the weight table and the draw records:
// both in LDS; lean (the table read from global memory through the caches,
// the records exchanged with ds_bpermute, so neither takes LDS); or the table
// in global memory and the records in LDS.
enum FwdMode : int { kFwdLds = 0, kFwdLean = 1, kFwdTableGlobal = 2 };

// A compiled forward kernel for some G: its block size and mode.
struct FwdInstance {
    uint32_t block;
    int mode;
};
// Instances compiled for G lanes per path (ascending block size); returns the
// count.
int forward_instances(uint32_t G, FwdInstance *out, int cap);

// Sub-chain window slots per lane for G lanes per path (diff <= 25).
inline uint32_t forward_window(uint32_t G) { return (25u + G - 1u) / G; }

// Row index of the per-lane zero row: at least S (past the segments) and at
// least window-1, so the redirected addresses scr - w*SLOT stay >= 0.
inline uint32_t forward_scratch_row(uint32_t G, uint32_t N)
{
    const uint32_t S = (N + G - 1) / G, w = forward_window(G);
    return S > w - 1 ? S : w - 1;
}

// Dynamic LDS bytes of the forward kernel: segment rows + zero row, table
// (kFwdLds only), weight-histogram bins, draw records (G > 1, not kFwdLean);
// at least the block reduction's scratch.
inline uint32_t forward_lds_bytes(uint32_t G, uint32_t block, uint32_t N, uint32_t tableStride,
      uint32_t histStride = 0, int mode = kFwdLds)
{
    const uint64_t rows = (uint64_t)forward_scratch_row(G, N) + 1;
    const uint64_t floats = rows * 3ull * block + (mode == kFwdLds ? (uint64_t)tableStride : 0ull) + histStride +
          (G > 1 && mode != kFwdLean ? 4ull * block : 0ull);
    const uint64_t reduceBytes = (uint64_t)((block + 63) / 64) * sizeof(LseAcc);
    const uint64_t bytes = floats * 4ull;
    return (uint32_t)(bytes > reduceBytes ? bytes : reduceBytes);
}

hipError_t launch_forward(const FwdParams &p, uint32_t G, uint32_t grid, uint32_t block, int mode, uint32_t lds,
      hipStream_t s);
hipError_t launch_finalize(const BlockPartial *parts, uint32_t n, double absorbLog2PerPath, FinalOut *out,
      hipStream_t s);

// Exit-direction histogram (histogram.hip).  Pass 1 bins the scaled path
// weights 2^(log2 w - M) of a chunk of paths into an LDS tile of bins and
// writes the tile to partial[chunk][bin]; pass 2 sums the chunks per bin with
// compensated FP32 and divides by S in FP64.
struct HistShape {
    uint32_t numBins, tileBins, numTiles, numChunks;
    uint64_t chunkPaths;
};
HistShape hist_shape(uint64_t numPaths, uint32_t numBins);
hipError_t launch_histogram(const float2 *pathLog2, const uint32_t *pathBin, uint64_t numPaths,
      const FinalOut *fin, const HistShape &h, float *partial, double *hist, hipStream_t s);

}  // namespace twisty_hip
