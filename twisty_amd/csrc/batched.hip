// batched.hip -- persistent kernel over many heterogeneous experiments (gfx950).
//
// The reference runs one experiment per ExperimentRunner call, so a driver
// with many geometries (BeamSpreadExperiment.cpp:209-284) launches them one
// after another.  Here one launch covers a whole batch: the experiments' paths
// are cut into work items (a range of paths of one experiment), the items are
// ordered longest first, and a grid of a few blocks per CU pulls them from a
// global counter until the list is empty, so short and long experiments share
// the machine without a tail of idle CUs.  Each item runs the same per-path
// code as the plain kernel (perturb_path / weigh_state) with the lanes-per-path
// G of its experiment, chosen from its segment count.
//
// MODE (FwdMode) places the tables and the draw records as in forward_kernel:
// kFwdLds keeps the experiment's table in LDS and reloads it only when the next
// item belongs to an experiment with another table; kFwdTableGlobal reads the
// tables from global memory (no reload, no barrier per switch) and keeps the
// draw records in LDS, so more paths stay resident per CU; kFwdLean also
// passes the draw records with ds_bpermute.  Each item leaves one LSE
// partial; a second kernel merges an experiment's partials in FP64.
#include "batched.hpp"
#include "path_kernels.hpp"

namespace twisty_hip {

template<int G, int BLOCK, int MODE>
__device__ __forceinline__ void run_item(const BatchParams &p, const BatchExp &e, const BatchItem &item,
      float *smem, const TableView &tv, LseAcc &acc, uint32_t &clamped)
{
    constexpr uint32_t SLOT = 3u * BLOCK;
    constexpr int WMAX = (25 + G - 1) / G;
    constexpr uint32_t groupsPerBlock = BLOCK / G;
    const uint32_t lane = threadIdx.x;
    const uint32_t j = lane & (G - 1);
    const uint32_t groupBase = lane - j;
    constexpr bool LEAN = MODE == kFwdLean;
    uint32_t *drawbuf = LEAN ? nullptr
                             : reinterpret_cast<uint32_t *>(smem + p.rows * SLOT + (MODE == kFwdLds ? p.tableStride : 0u));
    const uint32_t scr = e.scratchRow * SLOT + lane;
    smem[scr] = 0.0f;
    smem[scr + BLOCK] = 0.0f;
    smem[scr + 2 * BLOCK] = 0.0f;
    const uint32_t ldsBase = __builtin_amdgcn_groupstaticsize();
    const uint32_t laneB = ldsBase + lane * 4u;
    uint32_t scrB[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) scrB[w] = ldsBase + (scr - (uint32_t)w * SLOT) * 4u;
    const float *init = p.segs + e.segOffset;
    for (uint32_t q = lane / G; q < item.count; q += groupsPerBlock) {
        const uint64_t chain = item.begin + q;
        smem[scr] = 0.0f;  // cleared per chain, as in forward_kernel
        smem[scr + BLOCK] = 0.0f;
        smem[scr + 2 * BLOCK] = 0.0f;
        load_state<G, BLOCK>(smem, init, 3u, 1u, e.S, e.N, lane, j);
        // burn-in: the draws of twisty_hip_run's burn-in at pathOffset 0
        for (uint32_t b = 0; b < e.burnIn; ++b)
            perturb_path<G, BLOCK, LEAN>(laneB, lane, j, groupBase, scrB, drawbuf, e.seed, 0u,
                  (1ull << 63) | (chain * e.burnIn + b), e.K, e.N, e.maxDiff);
        for (uint32_t sIdx = 0; sIdx < e.chainLength; ++sIdx) {
            const uint64_t local = chain * e.chainLength + sIdx;
            if (local >= e.numPaths) break;  // the last chain may be shorter
            perturb_path<G, BLOCK, LEAN>(laneB, lane, j, groupBase, scrB, drawbuf, e.seed, 0u, local, e.K, e.N,
                  e.maxDiff);
            wave_sync_lds();
            const PathLog pl = weigh_state<G, BLOCK>(smem, lane, j, groupBase, e.N, e.invDs, tv, clamped);
            wave_sync_lds();  // joints read other lanes' columns before the next reset
            if (j == 0) {
                const bool invalid = (pl.flags & 1u) || ((pl.flags & 6u) == 6u);
                float outv;
                if (invalid) {
                    ++acc.nInvalid;
                    outv = __int_as_float(0x7fc00000);
                } else {
                    ++acc.nValid;
                    if (pl.flags & 2u) {
                        outv = -INFINITY;
                    } else if (pl.flags & 4u) {
                        outv = INFINITY;
                        lse_add(acc, INFINITY, 0.0f);
                    } else {
                        lse_add(acc, pl.lsum.hi, pl.lsum.lo);
                        const float l2 = 0.30102999566398119521f;
                        outv = fmaf(pl.lsum.hi, l2, fmaf(pl.lsum.lo, l2, e.absorbLog10PerPath));
                    }
                }
                if (p.perPathLog10) p.perPathLog10[e.pathOut + local] = outv;
            }
        }
    }
}

// Three waves per SIMD (168 VGPRs, no spills) for the modes that keep the draw
// records in LDS, so three 256-lane blocks share a CU wherever their LDS fits
// (short paths, or the tables in global memory).
template<int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK, MODE == kFwdLean ? 1 : 3) batched_kernel(const BatchParams p)
{
    // No static LDS: the byte addressing of perturb_one takes the dynamic
    // block to start at groupstaticsize(), so the item slot lives at its end.
    extern __shared__ float smem[];
    const uint32_t tabOff = p.rows * 3u * BLOCK;
    constexpr bool TABLE_LDS = MODE == kFwdLds;
    const uint32_t itemOff = (TABLE_LDS ? p.tableStride : 0u) + (MODE == kFwdLean ? 0u : 4u * BLOCK);
    uint32_t &sItem = *reinterpret_cast<uint32_t *>(smem + tabOff + itemOff);
    uint32_t curTable = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(p.queue, 1u);
        __syncthreads();
        const uint32_t it = sItem;
        if (it >= p.numItems) break;
        const BatchItem item = p.items[it];
        const BatchExp e = p.exps[item.exp];
        const BatchTable bt = p.tables[e.table];
        if (TABLE_LDS && e.table != curTable) {
            // The previous item ended with a block barrier, so no lane still
            // reads the old table.
            for (int i = threadIdx.x; i < bt.size; i += BLOCK) smem[tabOff + i] = p.tableData[bt.offset + i];
            __syncthreads();
            curTable = e.table;
        }
        const TableView tv = make_table_view(TABLE_LDS ? smem + tabOff : p.tableData + bt.offset, bt.size, bt.minK,
              bt.maxK, bt.step);
        LseAcc acc;
        lse_init(acc);
        uint32_t clamped = 0;
        switch (e.G) {
        case 2: run_item<2, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        case 4: run_item<4, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        case 8: run_item<8, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        case 16: run_item<16, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        case 32: run_item<32, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        default: run_item<64, BLOCK, MODE>(p, e, item, smem, tv, acc, clamped); break;
        }
        acc.nClamped = clamped;
        block_reduce_lse(acc, smem, p.itemPartials + it);
        // Every thread has read sItem before thread 0 rewrites it, and thread
        // 0's end of the reduction and its fetch of the next item are kept
        // apart by a barrier: with nothing between them the compiler threads
        // thread 0 from one to the other, and the structured control flow it
        // builds from that runs the loop's barriers for the other lanes of
        // wave 0 without thread 0, so they re-read the same item without end
        // (seen as a stall of the table-in-global-memory forms).
        __syncthreads();
    }
}

__global__ void batch_reset_kernel(uint32_t *queue) { *queue = 0u; }

__global__ void __launch_bounds__(256) batch_finalize_kernel(const BlockPartial *parts, const uint32_t *expItemStart,
      const uint32_t *expItems, const double *absorbLog2, FinalOut *outs)
{
    const uint32_t e = blockIdx.x;
    const uint32_t b = expItemStart[e];
    finalize_partials(parts, expItems + b, expItemStart[e + 1] - b, absorbLog2[e], outs + e);
}

bool batched_block_supported(uint32_t block, int mode)
{
    if (mode == kFwdTableGlobal) return block == 256;
    return (mode == kFwdLds || mode == kFwdLean) && (block == 256 || block == 512);
}

#define TWISTY_BATCHED_INSTANCES(X) X(256, 0) X(512, 0) X(256, 1) X(512, 1) X(256, 2)

static const void *batched_fn(uint32_t block, int mode)
{
#define X(B, M) \
    if (block == B && mode == M) return (const void *)batched_kernel<B, M>;
    TWISTY_BATCHED_INSTANCES(X)
#undef X
    return nullptr;
}

hipError_t batched_prepare(uint32_t block, int mode, uint32_t lds)
{
    const void *fn = batched_fn(block, mode);
    if (!fn) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_batched(const BatchParams &p, uint32_t grid, uint32_t block, int mode, uint32_t lds,
      const uint32_t *expItemStart, const uint32_t *expItems, const double *absorbLog2, uint32_t numExps,
      FinalOut *outs, hipStream_t s)
{
    if (!batched_fn(block, mode)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, s, p.queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
#define X(B, M) \
    if (block == B && mode == M) hipLaunchKernelGGL((batched_kernel<B, M>), dim3(grid), dim3(B), lds, s, p);
    TWISTY_BATCHED_INSTANCES(X)
#undef X
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(numExps), dim3(256), 0, s, p.itemPartials, expItemStart, expItems,
          absorbLog2, outs);
    return hipGetLastError();
}

}  // namespace twisty_hip
