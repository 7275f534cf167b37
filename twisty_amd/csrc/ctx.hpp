// ctx.hpp -- the context behind twisty_hip_ctx and the host-side helpers of
// the C ABI sources (capi.hip, batch_capi.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/twisty_hip.h"
#include "adjoint.hpp"
#include "forward.hpp"

using namespace twisty_hip;

#include <vector>

struct twisty_hip_batch;

namespace twisty_hip {
// Frees a batch's device resources and detaches it from its context; the
// batch object itself stays until twisty_hip_batch_destroy (batch_capi.hip).
void batch_release(twisty_hip_batch *b);
}  // namespace twisty_hip

struct twisty_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    uint32_t numCUs = 0;
    uint32_t ldsPerBlock = 0;
    uint32_t ldsPerCU = 0;
    std::string lastError;

    float *dTable = nullptr;
    uint32_t tableSize = 0;
    float minK = 0, maxK = 0, step = 0, ds = 0, absorption = 0;
    bool haveTables = false;

    float *dInitSeg = nullptr;
    uint32_t N = 0;
    float arclength = 0;
    float endDir[3] = { 0, 0, 1 };
    float startPoint[3] = { 0, 0, 0 };  // first point of the initial curve
    float boundary[13] = {};             // BoundaryConditions_CudaSafe of set_curve
    bool keepFinalState = false;         // run() keeps every path's final segments
    uint32_t burnIn = 0;                 // run(): unweighted paths at the start of every chain
    uint64_t lastFinalPaths = 0;         // paths whose final segments dFinalSeg holds
    bool haveCurve = false;

    BlockPartial *dPartials = nullptr;
    uint64_t partialsCap = 0;
    FinalOut *dFinal = nullptr;
    // This device's address of the static scratch of twisty_hip_run_noalloc
    // (looked up once per device by twisty_hip_create; never freed).
    unsigned char *dScratch = nullptr;

    float *dPerPath = nullptr;
    uint64_t perPathCap = 0;
    float *dPerJoint = nullptr;
    uint64_t perJointCap = 0;

    // Exit-direction histogram.
    uint32_t histTheta = 0, histPhi = 0;
    float2 *dPathLog2 = nullptr;
    uint64_t pathLog2Cap = 0;
    uint32_t *dPathBin = nullptr;
    uint64_t pathBinCap = 0;
    float *dHistPartial = nullptr;
    uint64_t histPartialCap = 0;
    double *dHist = nullptr;
    uint64_t histCap = 0;
    uint64_t lastHistPaths = 0;
    uint32_t lastHistBins = 0;

    // Weight-magnitude histogram.
    uint32_t whBins = 0;
    double whMin = 0.0, whMax = 0.0, whNormLog10 = 0.0;
    unsigned long long *dWeightHist = nullptr;
    uint64_t weightHistCap = 0;
    uint32_t lastWeightBins = 0;

    // Adjoint.
    float *dFinalSeg = nullptr;
    float *dGenMin = nullptr;  // min log10 weight of the independent-path run
    float *dMultiTables = nullptr;  // multi-ds tables of run_generated_multi
    float *dMultiMeta = nullptr;
    uint64_t multiTablesCap = 0, multiMetaCap = 0;
    uint32_t multiBins = 0, multiCount = 0, multiSegments = 0;
    float multiAbsorption = 0.0f;   // the absorption coefficient the multi-ds meta was formed with
    float *dChainLog = nullptr;     // chains form: the chain's log10 weight after each proposal
    uint64_t chainLogCap = 0;
    // Adjoint roulette of the exit-direction sweep: 0 = every path; the
    // default 1 keeps about a tenth of the paths (the table gradient always
    // takes every path).
    float adjKappa = 1.0f;
    // Directions of the initial segments for per-result statistics of the
    // sweep's gradient (twisty_hip_set_adjoint_directions).
    float *dDirs = nullptr;
    uint64_t dirsCap = 0;
    uint32_t numDirs = 0, dirsN = 0;
    std::vector<float> dirsHost;  // the same directions, for the host-side projection
    float *dDirStats = nullptr;
    uint64_t dirStatsCap = 0;
    uint32_t lastDirs = 0;       // directions the last adjoint reported
    double lastDirGrad[kMaxDirs] = {}, lastDirRouletteStd[kMaxDirs] = {}, lastDirMcStd[kMaxDirs] = {};
    AdjPick *dPick = nullptr;
    uint64_t pickCap = 0;
    uint32_t *dPickCount = nullptr;
    uint64_t pickCountCap = 0;
    uint64_t lastPicked = 0;   // paths the last adjoint visited
    uint64_t genMinCap = 0;
    uint64_t finalSegCap = 0;
    float *dGradTPartial = nullptr;
    uint64_t gradTPartialCap = 0;
    float *dGradSPartial = nullptr;
    uint64_t gradSPartialCap = 0;
    double *dGradT = nullptr;
    uint64_t gradTCap = 0;
    double *dGradS = nullptr;
    uint64_t gradSCap = 0;

    uint32_t schedG = 0, schedBlock = 0, schedBpc = 0;
    uint32_t adjG = 0, adjBlock = 0, adjBpc = 0;
    float lastFwdMs = 0.0f, lastTotalMs = 0.0f, lastAdjMs = 0.0f;

    std::vector<twisty_hip_batch *> batches;  // live batches, released with the context
};

namespace twisty_hip {

inline int fail(twisty_hip_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->lastError = msg;
    return code;
}

inline int hipfail(twisty_hip_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, TWISTY_HIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                                                                        \
    do {                                                                                                          \
        hipError_t _e = (expr);                                                                                   \
        if (_e != hipSuccess) return hipfail((ctx), _e, #expr);                                                   \
    } while (0)

inline bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

inline uint32_t table_stride(uint32_t n) { return (n + 63u) & ~63u; }

template<class T>
inline int ensure_dev(twisty_hip_ctx *ctx, T *&ptr, uint64_t &cap, uint64_t n, const char *what)
{
    if (n <= cap && ptr) return TWISTY_HIP_OK;
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    if (e != hipSuccess) return hipfail(ctx, e, what);
    if ((e = hipMalloc(&ptr, n * sizeof(T))) != hipSuccess) return hipfail(ctx, e, what);
    cap = n;
    return TWISTY_HIP_OK;
}

}  // namespace twisty_hip
