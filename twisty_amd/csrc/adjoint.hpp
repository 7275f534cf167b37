// adjoint.hpp -- parameter block and launchers of the adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "forward.hpp"

namespace twisty_hip {

// A path the roulette keeps, and the factor its contribution is scaled by.
struct AdjPick {
    uint64_t local;
    float factor;
    uint32_t pad;
};

struct AdjParams {
    const float *finalSeg;    // forward's final segments: (3i + c)*numPaths + local
    const float2 *pathLog2;   // forward's per-path (hi, lo) log2 weights
    const FinalOut *fin;      // forward's normaliser (M, S)
    const float *table;       // tableSize floats
    float *gradTablePartial;  // [gridDim.x][tableSize]
    float *gradSegPartial;    // [gridDim.x][3N]; written by the sweep instances only
    uint64_t seed;
    uint64_t pathOffset;
    uint64_t numPaths;
    uint32_t N, K, maxDiff, stream;
    uint32_t S;           // slots per lane = ceil(N / G)
    uint32_t scratchRow;  // as forward_scratch_row(G, N)
    int32_t tableSize;
    uint32_t tableStride;
    uint32_t segStride;   // floats reserved for the 3N segment gradient (multiple of 64)
    float minK, maxK, step, invDs, invStep;
    const AdjPick *pick;         // nullable: visit only these paths (roulette)
    const uint32_t *pickCount;   // their number (written by adjoint_pick_kernel)
    uint32_t tableGrad;          // 1: scatter dE/dT (sweep instances may leave it to a table-only pass)
    // Directions of the initial segments (sweep instances): numDirs vectors of
    // 3N floats, dirs[d * 3N + 3i + c]; per visited path the kernel projects
    // the path's dE/d(initial segments) onto them and accumulates the sums of
    // adj_dir_stats into dirStats[block][kDirStatsMax].
    const float *dirs;
    uint32_t numDirs;
    float *dirStats;
};

// Per-path sums of the direction statistics, for direction d (x = the path's
// contribution w f dln w/d(dir) in scaled weight units, w its scaled weight,
// f = 1/q its roulette factor):
//   [3d] sum x^2 (1 - q), [3d + 1] sum x^2 q, [3d + 2] sum x w; [3 kMaxDirs] sum w^2 f.
constexpr uint32_t kMaxDirs = 4;
constexpr uint32_t kDirStatsMax = 3 * kMaxDirs + 1;

// Dynamic LDS bytes: segment rows (and adjoint rows for the sweep), table,
// table gradient, segment gradient and draw records (sweep, G > 1).
inline uint32_t adjoint_lds_bytes(uint32_t G, uint32_t block, uint32_t N, uint32_t tableStride, bool sweep)
{
    const uint64_t rows = (uint64_t)forward_scratch_row(G, N) + 1;
    const uint64_t segStride = ((3ull * N) + 63ull) & ~63ull;
    uint64_t floats = rows * 3ull * block * (sweep ? 2ull : 1ull) + 2ull * tableStride;
    if (sweep) floats += segStride + (G > 1 ? 4ull * block : 0ull);
    return (uint32_t)(floats * 4ull);
}

int adjoint_block_sizes(uint32_t G, bool sweep, uint32_t *out, int cap);
hipError_t launch_adjoint(const AdjParams &p, uint32_t G, bool sweep, uint32_t grid, uint32_t block, uint32_t lds,
      hipStream_t s);
// Russian roulette over the paths before the adjoint: a path of scaled weight
// w (omega = w / S) is kept with probability q = min(1, omega * numPaths / kappa)
// and its contribution scaled by 1 / q; kept paths are appended to pick
// (*count reset by the caller).  The draw is Philox4x32-10 with counter
// (0xA0000000, 0, path lo, path hi), key = seed.
hipError_t launch_adjoint_pick(const float2 *pathLog2, const FinalOut *fin, uint64_t numPaths, uint64_t pathOffset,
      uint64_t seed, float kappa, AdjPick *pick, uint32_t *count, hipStream_t s);
// out[e] = (sum over blocks of partial[block][e]) / (S ln 10), compensated FP32
// sums over blocks, FP64 scale.
hipError_t launch_grad_reduce(const float *partial, uint32_t numBlocks, uint32_t numEntries, const FinalOut *fin,
      double *out, hipStream_t s);

}  // namespace twisty_hip
