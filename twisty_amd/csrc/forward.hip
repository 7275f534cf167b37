// forward.hip -- plain forward kernel of the optimal-perturb hot path (gfx950).
//
// One group of G lanes owns one path (or one Markov chain of chainLength paths).
// The state is the N segment vectors d_i = p_{i+1} - p_i, not the points: a
// rigid rotation of points left+1..right-1 about the left point
// (FullExperimentRunnerOptimalPerturb.cpp:495-519) is exactly a rotation of
// d_left..d_{right-1}, and the rotation axis p_right - p_left is their sum.
// Segment vectors keep each curvature free of the 1/ds amplification that
// differencing rounded positions introduces; they are normalised into tangents
// only when the path is weighted.
//
// LDS layout per block (BLOCK lanes, S = ceil(N/G) slots per lane):
//   [state: S * 3 * BLOCK floats][table: tableStride floats][draws: 4*BLOCK words if G > 1]
//   state element (slot s, component c) of lane L is at s*3*BLOCK + c*BLOCK + L.
// Lane j of a group stores segments i = s*G + j, so every lane touches only its
// own column: bank = L mod 32 for any slot, i.e. the data-dependent sub-chain
// accesses are bank-conflict free, and sub-chain slot w of a perturbation is a
// compile-time offset from one per-perturbation base address.
#include "device_common.hpp"
#include "forward.hpp"
#include "path_kernels.hpp"

namespace twisty_hip {

// Exit-direction bin of the unit tangent t: cos(theta) = t.e over nTheta bins
// of [-1, 1], phi = atan2(t.f2, t.f1) over nPhi bins of [-pi, pi), clamped to
// the grid (build-defined; the reference bins only weight magnitudes,
// FullExperimentRunnerOptimalPerturb.cpp:26-37).
__device__ __forceinline__ uint32_t exit_bin(const FwdParams &p, float t0, float t1, float t2)
{
    const float ct = fmaf(t0, p.exitE[0], fmaf(t1, p.exitE[1], t2 * p.exitE[2]));
    const float c1 = fmaf(t0, p.exitF1[0], fmaf(t1, p.exitF1[1], t2 * p.exitF1[2]));
    const float c2 = fmaf(t0, p.exitF2[0], fmaf(t1, p.exitF2[1], t2 * p.exitF2[2]));
    const float phi = atan2f(c2, c1);
    int it = (int)floorf((ct + 1.0f) * 0.5f * (float)p.nTheta);
    int ip = (int)floorf((phi + 3.14159265358979f) * (0.5f / 3.14159265358979f) * (float)p.nPhi);
    it = min(max(it, 0), (int)p.nTheta - 1);
    ip = min(max(ip, 0), (int)p.nPhi - 1);
    return (uint32_t)it * p.nPhi + (uint32_t)ip;
}

// Bin of a valid path in the weight-magnitude histogram of
// FullExperimentRunnerOptimalPerturb.cpp:562-575: the weight
// 10^(log10 w + normaliser) in FP64, below the range -> bin 0, above -> the
// last bin, else floor((w - min) / (max - min) * bins) (clamped: the reference
// indexes one past the end for w == max).  Zero weight -> bin 0.
__device__ __forceinline__ uint32_t weight_bin(const FwdParams &p, uint32_t flags, FF lsum)
{
    if (flags & 2u) return 0u;
    if (flags & 4u) return p.weightBins - 1u;
    const double l10 = ((double)lsum.hi + (double)lsum.lo) * 0.30102999566398119521 + p.weightLog10Offset;
    const double w = exp10(l10);
    if (w < p.weightMin) return 0u;
    if (w > p.weightMax) return p.weightBins - 1u;
    const double b = (w - p.weightMin) / (p.weightMax - p.weightMin) * (double)p.weightBins;
    const uint32_t bi = (uint32_t)b;
    return bi < p.weightBins ? bi : p.weightBins - 1u;
}

// MODE (FwdMode): kFwdLds keeps the table and the draw records in LDS.  The
// other two read the table from global memory (through the caches; it is only
// touched by the weighting, once per path) so that more paths stay resident
// per CU; kFwdLean also passes the draw records with ds_bpermute instead of an
// LDS buffer, so the LDS holds only path state and histogram bins.
template<int G, int BLOCK, int MODE>
__global__ void __launch_bounds__(BLOCK) forward_kernel(const FwdParams p)
{
    constexpr bool TABLE_GLOBAL = MODE != kFwdLds;
    constexpr bool LEAN = MODE == kFwdLean;
    constexpr uint32_t SLOT = 3u * BLOCK;  // floats per slot row
    extern __shared__ float smem[];
    const uint32_t N = p.N, S = p.S, K = p.K;
    // Rows 0..S-1 hold segments; row p.scratchRow (>= S and >= WMAX-1, so that
    // scr - w*SLOT never underflows) is the lane's zero row.
    const uint32_t tabOff = (p.scratchRow + 1) * SLOT;
    const uint32_t histOff = tabOff + (TABLE_GLOBAL ? 0u : p.tableStride);
    const uint32_t drawOff = histOff + p.histStride;
    uint32_t *drawbuf = LEAN ? nullptr : reinterpret_cast<uint32_t *>(smem + drawOff);
    uint32_t *whist = reinterpret_cast<uint32_t *>(smem + histOff);

    if constexpr (!TABLE_GLOBAL)
        for (int i = threadIdx.x; i < p.tableSize; i += BLOCK) smem[tabOff + i] = p.table[i];
    if (p.weightHist)
        for (uint32_t i = threadIdx.x; i < p.weightBins; i += BLOCK) whist[i] = 0u;
    __syncthreads();
    const TableView tv = make_table_view(TABLE_GLOBAL ? p.table : smem + tabOff, p.tableSize, p.minK, p.maxK, p.step);

    const uint32_t lane = threadIdx.x;
    const uint32_t j = lane & (G - 1);
    const uint32_t groupBase = lane - j;
    constexpr uint32_t groupsPerBlock = BLOCK / G;
    const uint32_t scr = p.scratchRow * SLOT + lane;
    smem[scr] = 0.0f;
    smem[scr + BLOCK] = 0.0f;
    smem[scr + 2 * BLOCK] = 0.0f;
    const uint32_t ldsBase = __builtin_amdgcn_groupstaticsize();
    const uint32_t laneB = ldsBase + lane * 4u;
    constexpr int WMAX = (25 + G - 1) / G;
    uint32_t scrB[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) scrB[w] = ldsBase + (scr - (uint32_t)w * SLOT) * 4u;

    LseAcc acc;
    lse_init(acc);
    uint32_t clamped = 0;

    for (uint64_t chain = (uint64_t)blockIdx.x * groupsPerBlock + lane / G; chain < p.numChains;
          chain += (uint64_t)gridDim.x * groupsPerBlock) {
        for (uint32_t sIdx = 0; sIdx < p.chainLength; ++sIdx) {
            const uint64_t local = chain * p.chainLength + sIdx;
            if (local >= p.numPaths) break;  // the last chain may be shorter
            const uint64_t path = p.pathOffset + local;
            if (sIdx == 0) {
                // A path whose rotation turned NaN stores NaN * 0 into the
                // scratch row; clearing it per chain keeps that path's
                // neighbours in the lane independent of it.
                smem[scr] = 0.0f;
                smem[scr + BLOCK] = 0.0f;
                smem[scr + 2 * BLOCK] = 0.0f;
                load_state<G, BLOCK>(smem, p.initSeg, 3u, 1u, S, N, lane, j);
                // Burn-in (FullExperimentRunnerOptimalPerturb.cpp:559-561):
                // states perturbed, never weighted; draws from path ids with
                // the top bit set, which weighted paths never reach.
                const uint64_t burnBase = (p.pathOffset / p.chainLength + chain) * (uint64_t)p.burnIn;
                for (uint32_t b = 0; b < p.burnIn; ++b)
                    perturb_path<G, BLOCK, LEAN>(laneB, lane, j, groupBase, scrB, drawbuf, p.seed, p.stream,
                          (1ull << 63) | (burnBase + b), K, N, p.maxDiff);
            }

            perturb_path<G, BLOCK, LEAN>(laneB, lane, j, groupBase, scrB, drawbuf, p.seed, p.stream, path, K, N,
                  p.maxDiff);

            // Weighting: tangents (CurvePerturbUtils.cpp:147-154), curvature
            // (CurvePerturbUtils.cpp:219-243), lookup and log
            // (PathWeighters.cpp:41-96).  Lane j handles joints i = j mod G.
            if constexpr (G > 1) wave_sync_lds();
            if (p.finalSeg) {
                for (uint32_t s = 0; s < S; ++s) {
                    const uint32_t i = s * G + j;
                    if (i < N) {
                        float *o = p.finalSeg + (uint64_t)(3 * i) * p.numPaths + local;
                        o[0] = smem[s * SLOT + lane];
                        o[p.numPaths] = smem[s * SLOT + BLOCK + lane];
                        o[2 * p.numPaths] = smem[s * SLOT + 2 * BLOCK + lane];
                    }
                }
            }
            FF lsum { 0.0f, 0.0f };
            uint32_t flags = 0;  // 1 = NaN, 2 = -inf, 4 = +inf segment log
            if constexpr (G > 1) {
                // The shared weighting loop, with the per-joint outputs when
                // asked for (curvatures; exit-direction bin from t_{N-2}).
                PathLog pl;
                if (!p.perJointCurv && !p.pathBin) {
                    pl = weigh_state<G, BLOCK>(smem, lane, j, groupBase, N, p.invDs, tv, clamped);
                } else {
                    pl = weigh_state<G, BLOCK>(smem, lane, j, groupBase, N, p.invDs, tv, clamped,
                          [&](uint32_t i, const JointEval &je) {
                              if (p.perJointCurv) p.perJointCurv[local * (uint64_t)(N - 1) + i] = je.kappa;
                              if (p.pathBin && i == N - 2) p.pathBin[local] = exit_bin(p, je.t0, je.t1, je.t2);
                          });
                }
                lsum = pl.lsum;
                flags = pl.flags;
            } else {
                // One lane per path: each segment's tangent is formed once
                // and carried to the next joint.
                float ta0, ta1, ta2;
                unit_tangent(smem[lane], smem[BLOCK + lane], smem[2 * BLOCK + lane], ta0, ta1, ta2);
                for (uint32_t i = 0; i + 1 < N; ++i) {
                    const uint32_t bi = (i + 1) * SLOT + lane;
                    const float b0 = smem[bi], b1 = smem[bi + BLOCK], b2 = smem[bi + 2 * BLOCK];
                    float tb0, tb1, tb2;
                    const float ib = unit_tangent(b0, b1, b2, tb0, tb1, tb2);
                    const float kappa = joint_diff(ta0, ta1, ta2, b0, b1, b2, ib, p.invDs).kappa;
                    if (p.perJointCurv) p.perJointCurv[local * (uint64_t)(N - 1) + i] = kappa;
                    if (p.pathBin && i == N - 2) p.pathBin[local] = exit_bin(p, ta0, ta1, ta2);
                    add_joint_log(lsum, flags, log2f(lookup_weight(tv, kappa, clamped)), true);
                    ta0 = tb0;
                    ta1 = tb1;
                    ta2 = tb2;
                }
            }
            if constexpr (G > 1) wave_sync_lds();  // joints read other lanes' columns before the next reset
            if (j == 0) {
                const bool invalid = (flags & 1u) || ((flags & 6u) == 6u);
                float outv;
                if (invalid) {
                    ++acc.nInvalid;
                    outv = __int_as_float(0x7fc00000);
                } else {
                    ++acc.nValid;
                    if (flags & 2u) {
                        outv = -INFINITY;
                    } else if (flags & 4u) {
                        outv = INFINITY;
                        lse_add(acc, INFINITY, 0.0f);
                    } else {
                        lse_add(acc, lsum.hi, lsum.lo);
                        outv = fmaf(lsum.hi, p.log10of2, fmaf(lsum.lo, p.log10of2, p.absorbLog10PerPath));
                    }
                }
                if (p.perPathLog10) p.perPathLog10[local] = outv;
                if (p.weightHist && !invalid)
                    atomicAdd(&whist[weight_bin(p, flags, lsum)], 1u);
                if (p.pathLog2) {
                    float2 v;
                    if (invalid) v = make_float2(__int_as_float(0x7fc00000), 0.0f);
                    else if (flags & 6u) v = make_float2((flags & 2u) ? -INFINITY : INFINITY, 0.0f);
                    else v = make_float2(lsum.hi, lsum.lo);
                    p.pathLog2[local] = v;
                }
            }
        }
    }
    acc.nClamped = clamped;
    if (p.weightHist) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < p.weightBins; i += BLOCK)
            if (whist[i]) atomicAdd(&p.weightHist[i], (unsigned long long)whist[i]);
    }
    block_reduce_lse(acc, smem, p.partials + blockIdx.x);
}

// FP64 cross-block normaliser: combines the per-block partials in a fixed
// order (deterministic for a fixed grid) and adds the per-path absorption term.
__global__ void __launch_bounds__(256) finalize_kernel(const BlockPartial *parts, uint32_t n,
      double absorbLog2PerPath, FinalOut *out)
{
    finalize_partials(parts, nullptr, n, absorbLog2PerPath, out);
}

// ---------------------------------------------------------------------------
// Instances: (lanes per path, block size, FwdMode).  Sub-chain slot offsets
// are immediates, so BLOCK is a template parameter; the set below covers the
// shapes the schedule resolver chooses from.  A block size can have one
// instance per mode; the resolver tries the table-in-LDS one first.  For
// N = 64 (G = 4) a 640-lane block with the table in LDS measured slower than
// 512; the lean 768-lane block runs 12 waves per CU.  320 lanes (G = 32, 64)
// fit five waves where 384 no longer fits: N = 1024 486 -> 469 ms against 256
// lanes; 352 lanes (G = 32: eleven paths, the last wave half full) fit too,
// 433 -> 405 ms at N = 1024, K = 200.  With the table in global memory 384
// lanes (twelve paths) fit at N = 1024.
// ---------------------------------------------------------------------------
#define TWISTY_FWD_INSTANCES(X)                                                                                   \
    X(1, 128, 0) X(1, 192, 0) X(2, 192, 0) X(2, 256, 0) X(2, 320, 0) X(4, 256, 0) X(4, 384, 0) X(4, 512, 0)       \
    X(4, 768, 1) X(8, 512, 0) X(8, 768, 0) X(8, 1024, 0) X(16, 512, 0) X(16, 1024, 0) X(32, 256, 0)              \
    X(32, 320, 0) X(32, 352, 0) X(32, 384, 0) X(32, 384, 2) X(32, 512, 0) X(32, 1024, 0) X(64, 256, 0)           \
    X(64, 320, 0) X(64, 512, 0) X(64, 640, 0) X(64, 704, 0) X(64, 1024, 0)

template<int G, int B, int M>
static hipError_t launch_gb(const FwdParams &p, uint32_t grid, uint32_t lds, hipStream_t s)
{
    auto *fn = forward_kernel<G, B, M>;
    hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(B), lds, s, p);
    return hipGetLastError();
}

int forward_instances(uint32_t G, FwdInstance *out, int cap)
{
    int n = 0;
#define X(g, b, m)                                                                                                \
    if (G == g && n < cap) out[n++] = FwdInstance { b, m };
    TWISTY_FWD_INSTANCES(X)
#undef X
    return n;
}

hipError_t launch_forward(const FwdParams &p, uint32_t G, uint32_t grid, uint32_t block, int mode, uint32_t lds,
      hipStream_t s)
{
#define X(g, b, m)                                                                                                \
    if (G == g && block == b && mode == m) return launch_gb<g, b, m>(p, grid, lds, s);
    TWISTY_FWD_INSTANCES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_finalize(const BlockPartial *parts, uint32_t n, double absorbLog2PerPath, FinalOut *out,
      hipStream_t s)
{
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, s, parts, n, absorbLog2PerPath, out);
    return hipGetLastError();
}

}  // namespace twisty_hip
