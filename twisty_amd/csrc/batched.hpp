// batched.hpp -- device descriptors of the persistent batched kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "forward.hpp"

namespace twisty_hip {

// One experiment as the batched kernel reads it.
struct BatchExp {
    uint64_t numPaths;
    uint64_t seed;
    uint64_t pathOut;     // offset of its path 0 in the per-path output
    uint32_t N, K, maxDiff, G;
    uint32_t S;           // slots per lane = ceil(N / G)
    uint32_t scratchRow;  // forward_scratch_row(G, N)
    uint32_t segOffset;   // float offset of its N initial segment vectors
    uint32_t table;       // index into the table descriptors
    float invDs;          // N / arclength
    float absorbLog10PerPath;
    uint32_t chainLength; // paths per chain (>= 1)
    uint32_t burnIn;      // unweighted paths opening every chain
};

struct BatchTable {
    uint32_t offset;  // float offset of the entries
    int32_t size;
    float minK, maxK, step;
};

// A work item: chains [begin, begin + count) of one experiment (paths when
// its chain length is 1).
struct BatchItem {
    uint64_t begin;
    uint32_t exp;
    uint32_t count;
};

struct BatchParams {
    const BatchExp *exps;
    const BatchItem *items;
    const BatchTable *tables;
    const float *tableData;
    const float *segs;
    BlockPartial *itemPartials;  // one per item
    float *perPathLog10;         // nullable
    uint32_t *queue;             // next item; zeroed by the reset kernel
    uint32_t numItems;
    uint32_t rows;               // state rows reserved per lane (max scratchRow + 1)
    uint32_t tableStride;        // LDS floats reserved for a table
    uint32_t pad;
};

// LDS bytes: state rows, one table (kFwdLds), the draw records (not
// kFwdLean), the work-item slot.
inline uint32_t batched_lds_bytes(uint32_t block, uint32_t rows, uint32_t tableStride, int mode = kFwdLds)
{
    const uint64_t extra = (mode == kFwdLds ? (uint64_t)tableStride : 0ull) + (mode != kFwdLean ? 4ull * block : 0ull);
    return (uint32_t)(((uint64_t)rows * 3ull * block + extra + 4ull) * 4ull);
}

// Block sizes compiled per mode (FwdMode): 256 and 512 for kFwdLds and
// kFwdLean, 256 for kFwdTableGlobal (at 168 VGPRs a SIMD holds three waves,
// and a second block of 320 or 384 lanes did not become resident beside the
// first: those sizes ran at one block per CU).
bool batched_block_supported(uint32_t block, int mode);
// Sets the kernel's dynamic-LDS limit; call outside graph capture.
hipError_t batched_prepare(uint32_t block, int mode, uint32_t lds);
// Reset of the queue counter, the persistent kernel and the per-experiment
// FP64 finalize, all on stream s (capturable).
hipError_t launch_batched(const BatchParams &p, uint32_t grid, uint32_t block, int mode, uint32_t lds,
      const uint32_t *expItemStart, const uint32_t *expItems, const double *absorbLog2, uint32_t numExps,
      FinalOut *outs, hipStream_t s);

}  // namespace twisty_hip
