// generate.hpp -- parameters of the independent-path estimator kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "forward.hpp"

namespace twisty_hip {

struct GenParams {
    const float *table;     // tableSize floats (device)
    BlockPartial *partials;  // one per block
    float *perPathLog10;     // nullable
    float *perJointCurv;     // nullable: numPaths x (N - 1) joint curvatures
    float *minLog10;         // one float, +inf before the launch
    uint64_t seed;
    uint64_t pathOffset;
    uint64_t numPaths;
    float start[3], startDir[3], end[3], endDir[3];
    float ds;                // arclength / N in float (the generator's)
    float invDs;             // N / arclength (curvature)
    uint32_t N;
    uint32_t levels;         // halvings of the N - 2 free segments
    uint32_t leafSpan;       // 2 or 3
    uint32_t rows;           // point rows per lane = ceil((N + 1) / G)
    int32_t tableSize;
    uint32_t tableStride;
    float minK, maxK, step;
    float log10of2;
    float absorbLog10PerPath;
    // Multi-ds form (ExperimentBase.cpp:1388-1597): ds ~ U[dsLo, dsHi) per path
    // and the table whose ds is nearest (CachedMultiArclengthWeightLookupTable::
    // GetWeightLookupTable, PathWeightUtils.cpp:223-236), read from HBM.
    const float *multiTables;  // numTables x tableStride floats
    const float *multiMeta;    // per table: ds, minK, maxK, step, absorption log2 per path
    uint32_t numTables;
    float dsLo, dsHi;
    // Chains form (ExperimentBase.cpp:1599-1818): numChains chains of
    // chainLength proposals, the first chain pathOffset; ds = dsLo + x with x
    // exponential of rate extinction truncated to [0, maxDs] (oneMinusQ = 1 -
    // exp(-extinction maxDs)); normLog10 enters the acceptance test.
    uint64_t numChains;
    uint64_t chainLength;
    double oneMinusQ;
    double normLog10;
    float extinction;
    float *chainLog10;         // nullable: the chain's log10 weight after each proposal
};

// Forms of the independent-path estimator.
enum GenMode : int { kGenSingle = 0, kGenMulti = 1, kGenChains = 2 };

// Levels of the recursion (N - 2 halved until 2 or 3), -1 when N is not
// 2 + 2^k * {2, 3} with k >= 1.
inline int generate_levels(uint32_t N, uint32_t &leafSpan)
{
    if (N < 6) return -1;
    uint32_t span = N - 2;
    int k = 0;
    while (span > 3) {
        if (span & 1u) return -1;
        span >>= 1;
        ++k;
    }
    if (k < 1) return -1;
    leafSpan = span;
    return k;
}

// Lanes per path: the smallest power of two >= 2 with at most 16 point rows.
inline uint32_t generate_lanes(uint32_t N)
{
    uint32_t G = 2;
    while (G < 64 && (N + 1 + G - 1) / G > 16) G <<= 1;
    return G;
}

// tableStride 0 for the multi-ds form (its tables stay in HBM).
inline uint32_t generate_lds_bytes(uint32_t G, uint32_t block, uint32_t N, uint32_t tableStride)
{
    const uint64_t rows = (N + 1 + G - 1) / G;
    const uint64_t bytes = (rows * 3ull * block + tableStride) * 4ull;
    const uint64_t reduceBytes = (uint64_t)(block / 64) * sizeof(LseAcc);
    return (uint32_t)(bytes > reduceBytes ? bytes : reduceBytes);
}

hipError_t launch_generate(const GenParams &p, uint32_t G, int mode, uint32_t grid, uint32_t block, uint32_t lds,
      hipStream_t s);

}  // namespace twisty_hip
