// capi.hip -- C ABI of libtwisty_hip.so (declared in include/twisty_hip.h).
//
// Owns device memory, the stream and timing events of a context; validates
// every shape on the host before a kernel is launched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/twisty_hip.h"
#include "adjoint.hpp"
#include "ctx.hpp"
#include "generate.hpp"
#include "forward.hpp"

using namespace twisty_hip;

namespace {

// Static scratch of twisty_hip_run_noalloc, the entry point for callers that
// cannot allocate: the FP64 result in the first kScratchFinalBytes, then at
// most kNoAllocMaxBlocks per-block partials (every block of a launch writes
// its own partial and the finalize kernel reads exactly those, so nothing has
// to start at zero).  Every GPU the library runs on holds its own copy until
// the process ends; calls that overlap on one GPU share it and race.
constexpr uint32_t kNoAllocMaxBlocks = 4096;
constexpr size_t kScratchFinalBytes = 256;
constexpr size_t kScratchBytes = kScratchFinalBytes + (size_t)kNoAllocMaxBlocks * sizeof(BlockPartial);
static_assert(sizeof(FinalOut) <= kScratchFinalBytes, "the FP64 result must fit its slot");
constexpr int kScratchDevices = 64;

}  // namespace

// In host code `scratch` names a host-side stand-in, not device memory: only
// the per-device address from hipGetSymbolAddress (scratch_address) is passed
// to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex gScratchMu;
unsigned char *gScratchAddr[kScratchDevices] = {};

// This device's address of `scratch`, looked up once per device; the device
// must be current.
hipError_t scratch_address(int device, unsigned char **out)
{
    *out = nullptr;
    if (device < 0 || device >= kScratchDevices) return hipSuccess;  // run_noalloc reports it
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (!gScratchAddr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        gScratchAddr[device] = static_cast<unsigned char *>(p);
    }
    *out = gScratchAddr[device];
    return hipSuccess;
}

// LDS words the forward kernel reserves for the weight-magnitude histogram.
uint32_t hist_stride(const twisty_hip_ctx *ctx) { return ctx->whBins ? ((ctx->whBins + 63u) & ~63u) : 0u; }

// Resolves a schedule for segment count N.  Zero fields are filled in: G from N
// (about 16 segments per lane), then the largest compiled block size with the
// table in LDS whose LDS fits one block (an instance that reads the table from
// global memory instead when it keeps more lanes resident per CU), then as
// many blocks per CU as the per-CU LDS holds (at most 2048 lanes per CU).  A
// given block size takes the first of its instances whose LDS fits (the table
// in LDS first).  mode receives the chosen instance's FwdMode.
int resolve_schedule(const twisty_hip_ctx *ctx, uint32_t N, uint32_t &G, uint32_t &block, uint32_t &bpc,
      uint32_t &lds, std::string &why, int &mode)
{
    const uint32_t ts = table_stride(ctx->tableSize ? ctx->tableSize : 4096);
    if (G == 0) {
        // About 16 segments per lane; from N = 512 on, 32 per lane (two paths
        // per wave measured faster than one at N = 1024).
        const uint32_t spl = N >= 512 ? 32u : 16u;
        G = 1;
        while (G < 64 && (N + G - 1) / G > spl) G <<= 1;
    }
    if (!is_pow2(G) || G > 64) {
        why = "lanesPerPath must be a power of two <= 64";
        return TWISTY_HIP_ERR_ARG;
    }
    FwdInstance inst[32];
    const int ns = forward_instances(G, inst, 32);
    auto ldsOf = [&](const FwdInstance &x) { return forward_lds_bytes(G, x.block, N, ts, hist_stride(ctx), x.mode); };
    auto lanesPerCU = [&](const FwdInstance &x) {
        uint32_t b = ctx->ldsPerCU / ldsOf(x);
        if (b * x.block > 2048) b = 2048 / x.block;
        return (b < 1 ? 1u : b) * x.block;
    };
    int chosen = -1;
    if (block == 0) {
        for (int k = ns - 1; k >= 0 && chosen < 0; --k)
            if (inst[k].mode == kFwdLds && ldsOf(inst[k]) <= ctx->ldsPerBlock) chosen = k;
        for (int k = 0; k < ns; ++k)
            if (inst[k].mode != kFwdLds && ldsOf(inst[k]) <= ctx->ldsPerBlock &&
                  (chosen < 0 || lanesPerCU(inst[k]) > lanesPerCU(inst[chosen])))
                chosen = k;
        if (chosen < 0) {
            why = "no compiled block size fits the LDS of one block for N=" + std::to_string(N);
            return TWISTY_HIP_ERR_UNSUPPORTED;
        }
        block = inst[chosen].block;
    } else {
        int any = -1;
        for (int k = 0; k < ns; ++k) {
            if (inst[k].block != block) continue;
            if (any < 0) any = k;
            if (chosen < 0 && ldsOf(inst[k]) <= ctx->ldsPerBlock) chosen = k;
        }
        if (any < 0) {
            why = "no kernel instance for lanesPerPath=" + std::to_string(G) + " blockSize=" + std::to_string(block);
            return TWISTY_HIP_ERR_UNSUPPORTED;
        }
        if (chosen < 0) chosen = any;  // reported below as an LDS overflow
    }
    mode = inst[chosen].mode;
    lds = ldsOf(inst[chosen]);
    if (lds > ctx->ldsPerBlock) {
        why = "schedule needs " + std::to_string(lds) + " LDS bytes per block, limit " +
              std::to_string(ctx->ldsPerBlock);
        return TWISTY_HIP_ERR_UNSUPPORTED;
    }
    if (bpc == 0) {
        bpc = ctx->ldsPerCU / lds;
        if (bpc * block > 2048) bpc = 2048 / block;
        if (bpc < 1) bpc = 1;
    }
    return TWISTY_HIP_OK;
}

// End frame of the exit-direction bins: e = normalised end direction,
// f1 = the coordinate axis least aligned with e made orthogonal to it,
// f2 = e x f1 (FP64, rounded once).
void end_frame(const float endDir[3], float e[3], float f1[3], float f2[3])
{
    double E[3] = { endDir[0], endDir[1], endDir[2] };
    const double en = std::sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
    for (double &v : E) v /= en;
    const double ax = std::fabs(E[0]), ay = std::fabs(E[1]), az = std::fabs(E[2]);
    int idx = 0;
    if (ay < ax) idx = 1;
    if (az < std::min(ax, ay)) idx = 2;
    double A[3] = { 0, 0, 0 };
    A[idx] = 1.0;
    const double d = A[0] * E[0] + A[1] * E[1] + A[2] * E[2];
    double F1[3] = { A[0] - d * E[0], A[1] - d * E[1], A[2] - d * E[2] };
    const double fn = std::sqrt(F1[0] * F1[0] + F1[1] * F1[1] + F1[2] * F1[2]);
    for (double &v : F1) v /= fn;
    const double F2[3] = { E[1] * F1[2] - E[2] * F1[1], E[2] * F1[0] - E[0] * F1[2], E[0] * F1[1] - E[1] * F1[0] };
    for (int a = 0; a < 3; ++a) {
        e[a] = (float)E[a];
        f1[a] = (float)F1[a];
        f2[a] = (float)F2[a];
    }
}

struct FwdRequest {
    uint64_t numPaths = 0, pathOffset = 0, seed = 0;
    uint32_t K = 0, chainLength = 1, burnIn = 0;
    bool perPath = false, perJoint = false, pathLog2 = false, pathBin = false, finalSeg = false;
    // When set, the per-block partials and the FP64 result go here (at most
    // maxBlocks partials) instead of the context's own buffers, and no buffer
    // is allocated: the request must ask for no per-path outputs.
    BlockPartial *partials = nullptr;
    FinalOut *final = nullptr;
    uint32_t maxBlocks = 0;
};

// Validates a run, sizes the buffers it needs and enqueues forward + finalize
// on the context's stream between events ev0 (start), ev1 (forward done) and
// ev2 (finalize done).  Nothing is synchronised.
int enqueue_forward(twisty_hip_ctx *ctx, const FwdRequest &r, const char *who)
{
    const std::string w(who);
    if (!ctx->haveTables || !ctx->haveCurve) return fail(ctx, TWISTY_HIP_ERR_STATE, w + ": set_tables and set_curve first");
    if (r.numPaths == 0 || r.K == 0 || r.chainLength == 0)
        return fail(ctx, TWISTY_HIP_ERR_ARG, w + ": numPaths, perturbationsPerPath and chainLength must be >= 1");
    if (r.pathOffset % r.chainLength != 0)
        return fail(ctx, TWISTY_HIP_ERR_ARG, w + ": pathOffset must be a multiple of chainLength");
    if (r.numPaths > (1ull << 40)) return fail(ctx, TWISTY_HIP_ERR_ARG, w + ": numPaths > 2^40");
    const uint32_t N = ctx->N;
    uint32_t G = ctx->schedG, B = ctx->schedBlock, bpc = ctx->schedBpc, lds = 0;
    int mode = kFwdLds;
    std::string why;
    int rc = resolve_schedule(ctx, N, G, B, bpc, lds, why, mode);
    if (rc) return fail(ctx, rc, w + ": " + why);
    const uint32_t grid = ctx->numCUs * bpc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (r.partials) {
        if (grid > r.maxBlocks)
            return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED, w + ": the schedule launches " + std::to_string(grid) +
                        " blocks, more than the " + std::to_string(r.maxBlocks) + " partials of the static scratch");
    } else if ((rc = ensure_dev(ctx, ctx->dPartials, ctx->partialsCap, grid, "partials"))) {
        return rc;
    }
    BlockPartial *partials = r.partials ? r.partials : ctx->dPartials;
    FinalOut *final = r.final ? r.final : ctx->dFinal;
    if (r.perPath && (rc = ensure_dev(ctx, ctx->dPerPath, ctx->perPathCap, r.numPaths, "per-path buffer"))) return rc;
    const uint64_t nJoint = r.numPaths * (uint64_t)(N - 1);
    if (r.perJoint && (rc = ensure_dev(ctx, ctx->dPerJoint, ctx->perJointCap, nJoint, "per-joint buffer"))) return rc;
    if (r.pathLog2 && (rc = ensure_dev(ctx, ctx->dPathLog2, ctx->pathLog2Cap, r.numPaths, "path log2 buffer")))
        return rc;
    if (r.pathBin && (rc = ensure_dev(ctx, ctx->dPathBin, ctx->pathBinCap, r.numPaths, "path bin buffer"))) return rc;
    if (r.finalSeg &&
          (rc = ensure_dev(ctx, ctx->dFinalSeg, ctx->finalSegCap, r.numPaths * 3ull * N, "final segment buffer")))
        return rc;

    // Per-segment absorption factor, as PathWeighters.cpp:37 computes it
    // (float argument -> float exp), applied to every one of the N-1 joints.
    const double absorb = (double)std::exp((float)(-ctx->absorption * ctx->ds));
    const double absorbLog2PerPath = (double)(N - 1) * std::log2(absorb);

    FwdParams p {};
    p.initSeg = ctx->dInitSeg;
    p.table = ctx->dTable;
    p.partials = partials;
    p.perPathLog10 = r.perPath ? ctx->dPerPath : nullptr;
    p.perJointCurv = r.perJoint ? ctx->dPerJoint : nullptr;
    p.seed = r.seed;
    p.pathOffset = r.pathOffset;
    p.numChains = (r.numPaths + r.chainLength - 1) / r.chainLength;
    p.numPaths = r.numPaths;
    p.chainLength = r.chainLength;
    p.burnIn = r.burnIn;
    p.N = N;
    p.K = r.K;
    p.maxDiff = (N - 2 < 25u) ? (N - 2) : 25u;
    p.stream = 0;
    p.S = (N + G - 1) / G;
    p.scratchRow = forward_scratch_row(G, N);
    p.tableSize = (int32_t)ctx->tableSize;
    p.tableStride = table_stride(ctx->tableSize);
    p.minK = ctx->minK;
    p.maxK = ctx->maxK;
    p.step = ctx->step;
    p.invDs = (float)N / ctx->arclength;  // CurvePerturbUtils.cpp:223
    p.log10of2 = 0.30102999566398119521f;
    p.absorbLog10PerPath = (float)(absorbLog2PerPath * 0.30102999566398119521);
    p.pathLog2 = r.pathLog2 ? ctx->dPathLog2 : nullptr;
    p.pathBin = r.pathBin ? ctx->dPathBin : nullptr;
    p.finalSeg = r.finalSeg ? ctx->dFinalSeg : nullptr;
    end_frame(ctx->endDir, p.exitE, p.exitF1, p.exitF2);
    p.nTheta = ctx->histTheta;
    p.nPhi = ctx->histPhi;
    if (ctx->whBins) {
        if ((rc = ensure_dev(ctx, ctx->dWeightHist, ctx->weightHistCap, ctx->whBins, "weight histogram"))) return rc;
        HIP_TRY(ctx, hipMemsetAsync(ctx->dWeightHist, 0, ctx->whBins * sizeof(unsigned long long), ctx->stream));
        p.weightHist = ctx->dWeightHist;
        p.weightBins = ctx->whBins;
        p.histStride = hist_stride(ctx);
        p.weightMin = ctx->whMin;
        p.weightMax = ctx->whMax;
        p.weightLog10Offset = absorbLog2PerPath * 0.30102999566398119521 + ctx->whNormLog10;
    }
    ctx->lastWeightBins = ctx->whBins;

    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipError_t e = launch_forward(p, G, grid, B, mode, lds, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "forward kernel launch");
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    e = launch_finalize(partials, grid, absorbLog2PerPath, final, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "finalize kernel launch");
    HIP_TRY(ctx, hipEventRecord(ctx->ev2, ctx->stream));
    return TWISTY_HIP_OK;
}

void copy_partials(const FinalOut &fo, uint64_t numPaths, twisty_hip_partials *out)
{
    out->log10_sum_w = fo.log10SumW;
    out->log10_sum_w2 = fo.log10SumW2;
    out->max_log10_w = fo.maxLog10W;
    out->num_paths = numPaths;
    out->num_valid = fo.nValid;
    out->num_invalid = fo.nInvalid;
    out->num_clamped = fo.nClamped;
}

// Adjoint schedule: G from N (about 8 segments per lane: the sweep keeps two
// rows per segment), then the largest compiled block that fits.
int resolve_adjoint_schedule(const twisty_hip_ctx *ctx, uint32_t N, bool sweep, uint32_t &G, uint32_t &block,
      uint32_t &bpc, uint32_t &lds, std::string &why)
{
    const uint32_t ts = table_stride(ctx->tableSize ? ctx->tableSize : 4096);
    if (G == 0) {
        G = 1;
        while (G < 64 && (N + G - 1) / G > (sweep ? 8u : 16u)) G <<= 1;
    }
    if (!is_pow2(G) || G > 64) {
        why = "adjoint lanesPerPath must be a power of two <= 64";
        return TWISTY_HIP_ERR_ARG;
    }
    uint32_t sizes[32];
    const int ns = adjoint_block_sizes(G, sweep, sizes, 32);
    if (block == 0) {
        for (int k = ns - 1; k >= 0; --k) {
            if (adjoint_lds_bytes(G, sizes[k], N, ts, sweep) <= ctx->ldsPerBlock) {
                block = sizes[k];
                break;
            }
        }
        if (block == 0) {
            why = "no compiled adjoint block size fits the LDS of one block for N=" + std::to_string(N);
            return TWISTY_HIP_ERR_UNSUPPORTED;
        }
    }
    bool compiled = false;
    for (int k = 0; k < ns; ++k) compiled |= (sizes[k] == block);
    if (!compiled) {
        why = "no adjoint kernel instance for lanesPerPath=" + std::to_string(G) + " blockSize=" + std::to_string(block);
        return TWISTY_HIP_ERR_UNSUPPORTED;
    }
    lds = adjoint_lds_bytes(G, block, N, ts, sweep);
    if (lds > ctx->ldsPerBlock) {
        why = "adjoint schedule needs " + std::to_string(lds) + " LDS bytes per block, limit " +
              std::to_string(ctx->ldsPerBlock);
        return TWISTY_HIP_ERR_UNSUPPORTED;
    }
    if (bpc == 0) {
        bpc = ctx->ldsPerCU / lds;
        if (bpc * block > 2048) bpc = 2048 / block;
        if (bpc < 1) bpc = 1;
    }
    return TWISTY_HIP_OK;
}

}  // namespace

extern "C" {

int twisty_hip_version(void) { return 100; }

int twisty_hip_device_count(int *count)
{
    if (!count) return TWISTY_HIP_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return TWISTY_HIP_OK;
}

int twisty_hip_create(int device, twisty_hip_ctx **out)
{
    if (!out) return TWISTY_HIP_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return TWISTY_HIP_ERR_ARG;
    twisty_hip_ctx *ctx = new twisty_hip_ctx();
    ctx->device = device;
    auto bail = [&](hipError_t e, const char *what) {
        std::fprintf(stderr, "twisty_hip_create: %s: %s\n", what, hipGetErrorString(e));
        twisty_hip_destroy(ctx);
        return TWISTY_HIP_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    int v = 0;
    if ((e = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
        return bail(e, "CU count");
    ctx->numCUs = (uint32_t)v;
    if ((e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device)) != hipSuccess)
        return bail(e, "LDS per block");
    ctx->ldsPerBlock = (uint32_t)v;
    if ((e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device)) != hipSuccess)
        return bail(e, "LDS per CU");
    ctx->ldsPerCU = (uint32_t)v;
    if (ctx->ldsPerCU < ctx->ldsPerBlock) ctx->ldsPerCU = ctx->ldsPerBlock;
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "stream");
    if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreate(&ctx->ev1)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreate(&ctx->ev2)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreate(&ctx->ev3)) != hipSuccess) return bail(e, "event");
    if ((e = hipMalloc(&ctx->dFinal, sizeof(FinalOut))) != hipSuccess) return bail(e, "final buffer");
    if ((e = scratch_address(device, &ctx->dScratch)) != hipSuccess) return bail(e, "static scratch address");
    *out = ctx;
    return TWISTY_HIP_OK;
}

int twisty_hip_destroy(twisty_hip_ctx *ctx)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    // Teardown is best effort: errors here have no caller to act on them.
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    while (!ctx->batches.empty()) batch_release(ctx->batches.back());
    (void)hipFree(ctx->dTable);
    (void)hipFree(ctx->dInitSeg);
    (void)hipFree(ctx->dPartials);
    (void)hipFree(ctx->dFinal);
    (void)hipFree(ctx->dPerPath);
    (void)hipFree(ctx->dPerJoint);
    (void)hipFree(ctx->dPathLog2);
    (void)hipFree(ctx->dPathBin);
    (void)hipFree(ctx->dHistPartial);
    (void)hipFree(ctx->dHist);
    (void)hipFree(ctx->dFinalSeg);
    (void)hipFree(ctx->dGenMin);
    (void)hipFree(ctx->dPick);
    (void)hipFree(ctx->dMultiTables);
    (void)hipFree(ctx->dMultiMeta);
    (void)hipFree(ctx->dChainLog);
    (void)hipFree(ctx->dPickCount);
    (void)hipFree(ctx->dGradTPartial);
    (void)hipFree(ctx->dGradSPartial);
    (void)hipFree(ctx->dGradT);
    (void)hipFree(ctx->dGradS);
    (void)hipFree(ctx->dDirs);
    (void)hipFree(ctx->dDirStats);
    (void)hipFree(ctx->dWeightHist);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TWISTY_HIP_OK;
}

const char *twisty_hip_last_error(const twisty_hip_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->lastError.c_str();
}

int twisty_hip_set_tables(twisty_hip_ctx *ctx, const float *tables, uint32_t numBins, uint32_t numTables,
      float minCurvature, float maxCurvature, float curvatureStepSize, float ds, float absorption)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!tables || numBins < 2 || numBins > 65536 || numTables < 1 || numTables > 16)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_tables: need 2 <= numBins <= 65536 and 1 <= numTables <= 16");
    if (!(curvatureStepSize > 0.0f) || !std::isfinite(curvatureStepSize) || !std::isfinite(minCurvature) ||
          !std::isfinite(maxCurvature) || !(maxCurvature >= minCurvature) || !std::isfinite(ds) ||
          !std::isfinite(absorption))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_tables: non-finite or inverted curvature range");
    // Tables sharing one grid combine exactly by summing entries: the lerp of
    // PathWeighters.cpp:81-82 is linear in the table values.
    std::vector<float> sum(numBins);
    for (uint32_t b = 0; b < numBins; ++b) {
        double acc = 0.0;
        for (uint32_t t = 0; t < numTables; ++t) {
            const float v = tables[(uint64_t)t * numBins + b];
            if (std::isinf(v)) return fail(ctx, TWISTY_HIP_ERR_ARG, "set_tables: table contains inf");
            acc += (double)v;
        }
        sum[b] = (float)acc;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->tableSize != numBins) {
        HIP_TRY(ctx, hipFree(ctx->dTable));
        ctx->dTable = nullptr;
        HIP_TRY(ctx, hipMalloc(&ctx->dTable, (size_t)numBins * sizeof(float)));
    }
    HIP_TRY(ctx, hipMemcpy(ctx->dTable, sum.data(), (size_t)numBins * sizeof(float), hipMemcpyHostToDevice));
    ctx->tableSize = numBins;
    ctx->minK = minCurvature;
    ctx->maxK = maxCurvature;
    ctx->step = curvatureStepSize;
    ctx->ds = ds;
    ctx->absorption = absorption;
    ctx->haveTables = true;
    return TWISTY_HIP_OK;
}

int twisty_hip_set_curve(twisty_hip_ctx *ctx, const float *positions, uint32_t numSegments, const float boundary[13])
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!positions || !boundary || numSegments < 4 || numSegments > 65535)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_curve: need 4 <= numSegments <= 65535");
    const float arclength = boundary[12];
    if (!(arclength > 0.0f) || !std::isfinite(arclength))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_curve: arclength must be positive");
    const uint32_t N = numSegments;
    // Segment vectors p_{i+1} - p_i, differenced in FP64 and rounded once.
    std::vector<float> tan((size_t)N * 3);
    for (uint32_t i = 0; i < N; ++i) {
        double d[3], n2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            d[a] = (double)positions[3 * (i + 1) + a] - (double)positions[3 * i + a];
            n2 += d[a] * d[a];
        }
        if (!(n2 > 0.0) || !std::isfinite(n2))
            return fail(ctx, TWISTY_HIP_ERR_ARG, "set_curve: zero-length or non-finite segment " + std::to_string(i));
        for (int a = 0; a < 3; ++a) tan[3 * i + a] = (float)d[a];
        if (!((double)tan[3 * i] * tan[3 * i] + (double)tan[3 * i + 1] * tan[3 * i + 1] +
                      (double)tan[3 * i + 2] * tan[3 * i + 2] >
                    1e-60))
            return fail(ctx, TWISTY_HIP_ERR_ARG, "set_curve: segment " + std::to_string(i) + " underflows in float");
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->N != N) {
        HIP_TRY(ctx, hipFree(ctx->dInitSeg));
        ctx->dInitSeg = nullptr;
        HIP_TRY(ctx, hipMalloc(&ctx->dInitSeg, tan.size() * sizeof(float)));
    }
    HIP_TRY(ctx, hipMemcpy(ctx->dInitSeg, tan.data(), tan.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->N = N;
    ctx->arclength = arclength;
    for (int a = 0; a < 3; ++a) ctx->endDir[a] = boundary[9 + a];
    for (int a = 0; a < 3; ++a) ctx->startPoint[a] = positions[a];
    for (int a = 0; a < 13; ++a) ctx->boundary[a] = boundary[a];
    ctx->haveCurve = true;
    return TWISTY_HIP_OK;
}

int twisty_hip_set_schedule(twisty_hip_ctx *ctx, uint32_t lanesPerPath, uint32_t blockSize, uint32_t blocksPerCU)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (lanesPerPath && (!is_pow2(lanesPerPath) || lanesPerPath > 64))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_schedule: lanesPerPath must be a power of two <= 64");
    if (blockSize && (blockSize % 32 != 0 || blockSize > 1024))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_schedule: blockSize must be a multiple of 32 <= 1024");
    if (blocksPerCU > 32) return fail(ctx, TWISTY_HIP_ERR_ARG, "set_schedule: blocksPerCU <= 32");
    ctx->schedG = lanesPerPath;
    ctx->schedBlock = blockSize;
    ctx->schedBpc = blocksPerCU;
    return TWISTY_HIP_OK;
}

int twisty_hip_get_schedule(twisty_hip_ctx *ctx, uint32_t *lanesPerPath, uint32_t *blockSize,
      uint32_t *blocksPerCU, uint32_t *ldsBytes)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!ctx->haveCurve) return fail(ctx, TWISTY_HIP_ERR_STATE, "get_schedule: set_curve first");
    uint32_t G = ctx->schedG, B = ctx->schedBlock, bpc = ctx->schedBpc, lds = 0;
    int mode = kFwdLds;
    std::string why;
    const int rc = resolve_schedule(ctx, ctx->N, G, B, bpc, lds, why, mode);
    if (rc) return fail(ctx, rc, "get_schedule: " + why);
    if (lanesPerPath) *lanesPerPath = G;
    if (blockSize) *blockSize = B;
    if (blocksPerCU) *blocksPerCU = bpc;
    if (ldsBytes) *ldsBytes = lds;
    return TWISTY_HIP_OK;
}

int twisty_hip_run(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint32_t perturbationsPerPath,
      uint32_t chainLength, uint64_t seed, twisty_hip_partials *out, float *perPathLog10,
      float *perJointCurvature)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!out) return fail(ctx, TWISTY_HIP_ERR_ARG, "run: out is NULL");
    const bool hist = ctx->histTheta > 0;
    FwdRequest r;
    r.numPaths = numPaths;
    r.pathOffset = pathOffset;
    r.seed = seed;
    r.K = perturbationsPerPath;
    r.chainLength = chainLength;
    r.perPath = perPathLog10 != nullptr;
    r.perJoint = perJointCurvature != nullptr;
    r.pathLog2 = hist;
    r.pathBin = hist;
    r.finalSeg = ctx->keepFinalState;
    r.burnIn = ctx->burnIn;
    int rc = enqueue_forward(ctx, r, "run");
    if (rc) return rc;
    if (hist) {
        const uint32_t nb = ctx->histTheta * ctx->histPhi;
        const HistShape h = hist_shape(numPaths, nb);
        if ((rc = ensure_dev(ctx, ctx->dHistPartial, ctx->histPartialCap, (uint64_t)h.numChunks * nb, "histogram partials")))
            return rc;
        if ((rc = ensure_dev(ctx, ctx->dHist, ctx->histCap, nb, "histogram"))) return rc;
        hipError_t e = launch_histogram(ctx->dPathLog2, ctx->dPathBin, numPaths, ctx->dFinal, h, ctx->dHistPartial,
              ctx->dHist, ctx->stream);
        if (e != hipSuccess) return hipfail(ctx, e, "histogram kernel launch");
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev3, ctx->stream));
    FinalOut fo;
    HIP_TRY(ctx, hipMemcpyAsync(&fo, ctx->dFinal, sizeof(FinalOut), hipMemcpyDeviceToHost, ctx->stream));
    if (perPathLog10)
        HIP_TRY(ctx, hipMemcpyAsync(perPathLog10, ctx->dPerPath, numPaths * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    if (perJointCurvature)
        HIP_TRY(ctx, hipMemcpyAsync(perJointCurvature, ctx->dPerJoint, numPaths * (uint64_t)(ctx->N - 1) * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastFwdMs, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastTotalMs, ctx->ev0, ctx->ev3));
    ctx->lastAdjMs = 0.0f;
    ctx->lastHistPaths = hist ? numPaths : 0;
    ctx->lastHistBins = hist ? ctx->histTheta * ctx->histPhi : 0;
    ctx->lastFinalPaths = ctx->keepFinalState ? numPaths : 0;
    copy_partials(fo, numPaths, out);
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

int twisty_hip_run_noalloc(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset,
      uint32_t perturbationsPerPath, uint32_t chainLength, uint64_t seed, twisty_hip_partials *out)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!out) return fail(ctx, TWISTY_HIP_ERR_ARG, "run_noalloc: out is NULL");
    if (!ctx->dScratch)
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED, "run_noalloc: no static scratch on device " +
                    std::to_string(ctx->device) + " (devices 0.." + std::to_string(kScratchDevices - 1) + ")");
    if (ctx->histTheta || ctx->whBins || ctx->keepFinalState)
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED, "run_noalloc: the exit-direction and weight histograms and "
                    "the kept final state use buffers the library allocates; switch them off");
    FwdRequest r;
    r.numPaths = numPaths;
    r.pathOffset = pathOffset;
    r.seed = seed;
    r.K = perturbationsPerPath;
    r.chainLength = chainLength;
    r.burnIn = ctx->burnIn;
    r.final = reinterpret_cast<FinalOut *>(ctx->dScratch);
    r.partials = reinterpret_cast<BlockPartial *>(ctx->dScratch + kScratchFinalBytes);
    r.maxBlocks = kNoAllocMaxBlocks;
    int rc = enqueue_forward(ctx, r, "run_noalloc");
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev3, ctx->stream));
    FinalOut fo;
    HIP_TRY(ctx, hipMemcpyAsync(&fo, r.final, sizeof(FinalOut), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastFwdMs, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastTotalMs, ctx->ev0, ctx->ev3));
    ctx->lastAdjMs = 0.0f;
    ctx->lastHistPaths = 0;
    ctx->lastHistBins = 0;
    ctx->lastFinalPaths = 0;
    copy_partials(fo, numPaths, out);
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

int twisty_hip_set_adjoint_roulette(twisty_hip_ctx *ctx, float kappa)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!(kappa >= 0.0f) || !std::isfinite(kappa))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_roulette: kappa must be finite and >= 0");
    ctx->adjKappa = kappa;
    return TWISTY_HIP_OK;
}

int twisty_hip_last_adjoint_paths(twisty_hip_ctx *ctx, uint64_t *pathsVisited)
{
    if (!ctx || !pathsVisited) return TWISTY_HIP_ERR_ARG;
    *pathsVisited = ctx->lastPicked;
    return TWISTY_HIP_OK;
}

// The chains form's extra arguments (kGenChains).
struct ChainArgs {
    uint64_t numChains = 0, chainLength = 0;
    float extinction = 0.0f;
    double normLog10 = 0.0;
    float *chainLog10 = nullptr;
};

static int run_generated_impl(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint64_t seed,
      int mode, float dsLo, float dsHi, const ChainArgs &ch, twisty_hip_partials *out, double *minLog10W,
      float *perPathLog10, float *perJointCurvature)
{
    const bool multi = mode != kGenSingle;
    if (!out) return fail(ctx, TWISTY_HIP_ERR_ARG, "run_generated: out is NULL");
    if (!ctx->haveCurve) return fail(ctx, TWISTY_HIP_ERR_STATE, "run_generated: set_curve first");
    if (!multi && !ctx->haveTables) return fail(ctx, TWISTY_HIP_ERR_STATE, "run_generated: set_tables first");
    if (multi && ctx->multiCount == 0)
        return fail(ctx, TWISTY_HIP_ERR_STATE, "run_generated_multi: set_multi_tables first");
    if (mode == kGenMulti && !(dsLo > 0.0f && dsHi >= dsLo && std::isfinite(dsHi)))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "run_generated_multi: need 0 < minDs <= maxDs");
    if (mode == kGenChains) {
        if (!(dsLo > 0.0f && dsHi > 0.0f && std::isfinite(dsHi) && ch.extinction > 0.0f &&
                  (double)ch.extinction * (double)dsHi < 700.0 && std::isfinite(ch.normLog10)))
            return fail(ctx, TWISTY_HIP_ERR_ARG,
                  "run_generated_chains: need minDs > 0, maxDs > 0, extinction > 0 with extinction * maxDs < 700 "
                  "and a finite normaliser");
        if (ch.extinction != ctx->multiAbsorption)
            return fail(ctx, TWISTY_HIP_ERR_ARG,
                  "run_generated_chains: set_multi_tables must be given absorption = extinction (absorption + "
                  "scatter, the loss of every joint)");
        if (ch.numChains == 0 || ch.chainLength == 0 || ch.chainLength > (1ull << 32) ||
              ch.numChains > (1ull << 40) / ch.chainLength || pathOffset > (1ull << 62) / ch.chainLength - ch.numChains)
            return fail(ctx, TWISTY_HIP_ERR_ARG,
                  "run_generated_chains: need 1 <= numChains * chainLength <= 2^40 and chain ids below 2^62 / "
                  "chainLength");
    }
    if (numPaths == 0 || numPaths > (1ull << 40))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "run_generated: need 1 <= numPaths <= 2^40");
    const uint32_t N = ctx->N;
    uint32_t leafSpan = 0;
    const int levels = generate_levels(N, leafSpan);
    if (levels < 0)
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED,
              "run_generated: numSegments must be 2 + 2^k * 2 or 2 + 2^k * 3 with k >= 1 (the recursion of "
              "ResolveEvenNumberOfSegments halves N - 2 down to 2 or 3)");
    const uint32_t G = generate_lanes(N);
    const uint32_t rows = (N + 1 + G - 1) / G;
    if (rows > 16) return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED, "run_generated: numSegments above 1022");
    const uint32_t B = 256;
    const uint32_t bins = multi ? ctx->multiBins : ctx->tableSize;
    const uint32_t lds = generate_lds_bytes(G, B, N, multi ? 0u : table_stride(bins));
    if (lds > ctx->ldsPerBlock)
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED,
              "run_generated: needs " + std::to_string(lds) + " LDS bytes per block, limit " +
                    std::to_string(ctx->ldsPerBlock));
    uint32_t bpc = ctx->ldsPerCU / lds;
    if (bpc * B > 2048) bpc = 2048 / B;
    if (bpc < 1) bpc = 1;
    const uint32_t grid = ctx->numCUs * bpc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_dev(ctx, ctx->dPartials, ctx->partialsCap, grid, "partials"))) return rc;
    if (perPathLog10 && (rc = ensure_dev(ctx, ctx->dPerPath, ctx->perPathCap, numPaths, "per-path buffer"))) return rc;
    if (perJointCurvature &&
          (rc = ensure_dev(ctx, ctx->dPerJoint, ctx->perJointCap, numPaths * (uint64_t)(N - 1), "per-joint buffer")))
        return rc;
    if ((rc = ensure_dev(ctx, ctx->dGenMin, ctx->genMinCap, 1, "min buffer"))) return rc;
    if (ch.chainLog10 && (rc = ensure_dev(ctx, ctx->dChainLog, ctx->chainLogCap, numPaths, "chain buffer")))
        return rc;
    // the multi-ds form adds each path's own table absorption in the kernel
    const double absorbLog2PerPath =
          multi ? 0.0 : (double)(N - 1) * std::log2((double)std::exp((float)(-ctx->absorption * ctx->ds)));
    GenParams p {};
    p.table = ctx->dTable;
    p.partials = ctx->dPartials;
    p.perPathLog10 = perPathLog10 ? ctx->dPerPath : nullptr;
    p.perJointCurv = perJointCurvature ? ctx->dPerJoint : nullptr;
    p.minLog10 = ctx->dGenMin;
    p.seed = seed;
    p.pathOffset = pathOffset;
    p.numPaths = numPaths;
    for (int a = 0; a < 3; ++a) {
        p.start[a] = ctx->boundary[a];
        p.startDir[a] = ctx->boundary[3 + a];
        p.end[a] = ctx->boundary[6 + a];
        p.endDir[a] = ctx->boundary[9 + a];
    }
    p.ds = ctx->arclength / (float)N;  // ExperimentBase.cpp:1253
    p.invDs = (float)N / ctx->arclength;
    p.N = N;
    p.levels = (uint32_t)levels;
    p.leafSpan = leafSpan;
    p.rows = rows;
    p.tableSize = (int32_t)bins;
    p.tableStride = table_stride(bins);
    if (multi) {
        p.multiTables = ctx->dMultiTables;
        p.multiMeta = ctx->dMultiMeta;
        p.numTables = ctx->multiCount;
        p.dsLo = dsLo;
        p.dsHi = dsHi;
    }
    if (mode == kGenChains) {
        p.numChains = ch.numChains;
        p.chainLength = ch.chainLength;
        p.extinction = ch.extinction;
        p.oneMinusQ = 1.0 - std::exp(-(double)ch.extinction * (double)dsHi);
        p.normLog10 = ch.normLog10;
        p.chainLog10 = ch.chainLog10 ? ctx->dChainLog : nullptr;
    }
    p.minK = ctx->minK;
    p.maxK = ctx->maxK;
    p.step = ctx->step;
    p.log10of2 = 0.30102999566398119521f;
    p.absorbLog10PerPath = (float)(absorbLog2PerPath * 0.30102999566398119521);
    const float inf = INFINITY;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->dGenMin, &inf, sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipError_t e = launch_generate(p, G, mode, grid, B, lds, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "generate kernel launch");
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    e = launch_finalize(ctx->dPartials, grid, absorbLog2PerPath, ctx->dFinal, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "finalize kernel launch");
    HIP_TRY(ctx, hipEventRecord(ctx->ev2, ctx->stream));
    FinalOut fo;
    float mn = INFINITY;
    HIP_TRY(ctx, hipMemcpyAsync(&fo, ctx->dFinal, sizeof(FinalOut), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(&mn, ctx->dGenMin, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (perPathLog10)
        HIP_TRY(ctx, hipMemcpyAsync(perPathLog10, ctx->dPerPath, numPaths * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    if (perJointCurvature)
        HIP_TRY(ctx, hipMemcpyAsync(perJointCurvature, ctx->dPerJoint, numPaths * (uint64_t)(N - 1) * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    if (ch.chainLog10)
        HIP_TRY(ctx, hipMemcpyAsync(ch.chainLog10, ctx->dChainLog, numPaths * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastFwdMs, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastTotalMs, ctx->ev0, ctx->ev2));
    ctx->lastAdjMs = 0.0f;
    ctx->lastHistPaths = 0;
    ctx->lastHistBins = 0;
    ctx->lastFinalPaths = 0;
    ctx->lastWeightBins = 0;
    copy_partials(fo, numPaths, out);
    if (minLog10W) *minLog10W = fo.nValid ? (double)mn : INFINITY;
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

int twisty_hip_run_generated(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint64_t seed,
      twisty_hip_partials *out, double *minLog10W, float *perPathLog10, float *perJointCurvature)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    return run_generated_impl(ctx, numPaths, pathOffset, seed, kGenSingle, 0.0f, 0.0f, ChainArgs {}, out, minLog10W,
          perPathLog10, perJointCurvature);
}

int twisty_hip_set_multi_tables(twisty_hip_ctx *ctx, const float *tables, uint32_t numBins, uint32_t numTables,
      const float *tableMeta, float absorption, uint32_t numSegments)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!tables || !tableMeta || numBins < 2 || numBins > 65536 || numTables < 1 || numTables > 4096 ||
          numSegments < 2)
        return fail(ctx, TWISTY_HIP_ERR_ARG,
              "set_multi_tables: need tables, meta, 2 <= numBins <= 65536, 1 <= numTables <= 4096, numSegments >= 2");
    const uint32_t ts = table_stride(numBins);
    std::vector<float> data((size_t)ts * numTables, 0.0f), meta((size_t)5 * numTables);
    for (uint32_t t = 0; t < numTables; ++t) {
        const float *m = tableMeta + 4 * t;  // ds, minK, maxK, step
        if (!(m[0] > 0.0f) || !(m[3] > 0.0f) || !std::isfinite(m[1]) || !std::isfinite(m[2]) || !(m[2] >= m[1]) ||
              (t > 0 && !(m[0] > tableMeta[4 * (t - 1)])))
            return fail(ctx, TWISTY_HIP_ERR_ARG, "set_multi_tables: table " + std::to_string(t) +
                                                         " has a bad grid or its ds does not increase");
        for (uint32_t b = 0; b < numBins; ++b) {
            const float v = tables[(uint64_t)t * numBins + b];
            if (std::isinf(v)) return fail(ctx, TWISTY_HIP_ERR_ARG, "set_multi_tables: table contains inf");
            data[(size_t)t * ts + b] = v;
        }
        for (int k = 0; k < 4; ++k) meta[5 * t + k] = m[k];
        // per path: (N - 1) joints of expf(-absorption * table ds) (PathWeighters.cpp:37)
        meta[5 * t + 4] = (float)((double)(numSegments - 1) * std::log2((double)std::exp((float)(-absorption * m[0]))));
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_dev(ctx, ctx->dMultiTables, ctx->multiTablesCap, data.size(), "multi-ds tables"))) return rc;
    if ((rc = ensure_dev(ctx, ctx->dMultiMeta, ctx->multiMetaCap, meta.size(), "multi-ds meta"))) return rc;
    HIP_TRY(ctx, hipMemcpy(ctx->dMultiTables, data.data(), data.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(ctx->dMultiMeta, meta.data(), meta.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->multiBins = numBins;
    ctx->multiCount = numTables;
    ctx->multiSegments = numSegments;
    ctx->multiAbsorption = absorption;
    return TWISTY_HIP_OK;
}

int twisty_hip_run_generated_multi(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint64_t seed,
      float minDs, float maxDs, twisty_hip_partials *out, double *minLog10W, float *perPathLog10,
      float *perJointCurvature)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (ctx->multiCount && ctx->haveCurve && ctx->multiSegments != ctx->N)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "run_generated_multi: set_multi_tables was given another numSegments");
    return run_generated_impl(ctx, numPaths, pathOffset, seed, kGenMulti, minDs, maxDs, ChainArgs {}, out, minLog10W,
          perPathLog10, perJointCurvature);
}

int twisty_hip_run_generated_chains(twisty_hip_ctx *ctx, uint64_t numChains, uint64_t chainLength,
      uint64_t chainOffset, uint64_t seed, float minDs, float maxDs, float extinction, double pathNormalizerLog10,
      twisty_hip_partials *out, double *minLog10W, float *perPathLog10, float *chainLog10, float *perJointCurvature)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (ctx->multiCount && ctx->haveCurve && ctx->multiSegments != ctx->N)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "run_generated_chains: set_multi_tables was given another numSegments");
    ChainArgs ch;
    ch.numChains = numChains;
    ch.chainLength = chainLength;
    ch.extinction = extinction;
    ch.normLog10 = pathNormalizerLog10;
    ch.chainLog10 = chainLog10;
    const uint64_t n = (numChains && chainLength && numChains <= (1ull << 40) / chainLength) ? numChains * chainLength
                                                                                               : 0;
    return run_generated_impl(ctx, n ? n : 1, chainOffset, seed, kGenChains, minDs, maxDs, ch, out, minLog10W,
          perPathLog10, perJointCurvature);
}

int twisty_hip_set_burn_in(twisty_hip_ctx *ctx, uint32_t pathsPerChain)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (pathsPerChain > (1u << 24))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_burn_in: at most 2^24 burn-in paths per chain");
    ctx->burnIn = pathsPerChain;
    return TWISTY_HIP_OK;
}

int twisty_hip_set_keep_final_state(twisty_hip_ctx *ctx, int enable)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    ctx->keepFinalState = enable != 0;
    return TWISTY_HIP_OK;
}

int twisty_hip_get_final_positions(twisty_hip_ctx *ctx, float *positions, uint64_t numPaths)
{
    if (!ctx || !positions) return TWISTY_HIP_ERR_ARG;
    if (ctx->lastFinalPaths == 0)
        return fail(ctx, TWISTY_HIP_ERR_STATE, "get_final_positions: the last run kept no final state");
    if (numPaths != ctx->lastFinalPaths)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "get_final_positions: numPaths must equal the last run's path count");
    const uint32_t N = ctx->N;
    std::vector<float> seg((size_t)3 * N * numPaths);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(seg.data(), ctx->dFinalSeg, seg.size() * sizeof(float), hipMemcpyDeviceToHost));
    // Points p_0 + sum of the first i segments, accumulated in FP64.
    for (uint64_t q = 0; q < numPaths; ++q) {
        double acc[3] = { ctx->startPoint[0], ctx->startPoint[1], ctx->startPoint[2] };
        float *out = positions + q * 3ull * (N + 1);
        for (int a = 0; a < 3; ++a) out[a] = ctx->startPoint[a];
        for (uint32_t i = 0; i < N; ++i) {
            for (int a = 0; a < 3; ++a) {
                acc[a] += (double)seg[(uint64_t)(3 * i + a) * numPaths + q];
                out[3 * (i + 1) + a] = (float)acc[a];
            }
        }
    }
    return TWISTY_HIP_OK;
}

int twisty_hip_set_weight_histogram(twisty_hip_ctx *ctx, uint32_t numBins, double minWeight, double maxWeight,
      double pathNormalizerLog10)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (numBins > 4096 || (numBins && !(maxWeight > minWeight && std::isfinite(minWeight) && std::isfinite(maxWeight) &&
                                 std::isfinite(pathNormalizerLog10))))
        return fail(ctx, TWISTY_HIP_ERR_ARG,
              "set_weight_histogram: need numBins <= 4096 and finite minWeight < maxWeight (0 bins = off)");
    ctx->whBins = numBins;
    ctx->whMin = minWeight;
    ctx->whMax = maxWeight;
    ctx->whNormLog10 = pathNormalizerLog10;
    return TWISTY_HIP_OK;
}

int twisty_hip_get_weight_histogram(twisty_hip_ctx *ctx, uint64_t *counts)
{
    if (!ctx || !counts) return TWISTY_HIP_ERR_ARG;
    if (ctx->lastWeightBins == 0)
        return fail(ctx, TWISTY_HIP_ERR_STATE, "get_weight_histogram: the last run had no weight histogram");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(counts, ctx->dWeightHist, ctx->lastWeightBins * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return TWISTY_HIP_OK;
}

int twisty_hip_set_exit_histogram(twisty_hip_ctx *ctx, uint32_t nTheta, uint32_t nPhi)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if ((nTheta == 0) != (nPhi == 0) || (uint64_t)nTheta * nPhi > (1u << 20))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_exit_histogram: need nTheta, nPhi both 0 (off) or nTheta*nPhi <= 2^20");
    ctx->histTheta = nTheta;
    ctx->histPhi = nPhi;
    return TWISTY_HIP_OK;
}

int twisty_hip_get_exit_histogram(twisty_hip_ctx *ctx, double *histogram, uint32_t *perPathBin, uint64_t numPaths)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (ctx->lastHistBins == 0)
        return fail(ctx, TWISTY_HIP_ERR_STATE, "get_exit_histogram: the last run had no exit histogram");
    if (perPathBin && numPaths != ctx->lastHistPaths)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "get_exit_histogram: numPaths must equal the last run's path count");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (histogram)
        HIP_TRY(ctx, hipMemcpy(histogram, ctx->dHist, ctx->lastHistBins * sizeof(double), hipMemcpyDeviceToHost));
    if (perPathBin)
        HIP_TRY(ctx, hipMemcpy(perPathBin, ctx->dPathBin, numPaths * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return TWISTY_HIP_OK;
}

int twisty_hip_set_adjoint_schedule(twisty_hip_ctx *ctx, uint32_t lanesPerPath, uint32_t blockSize,
      uint32_t blocksPerCU)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (lanesPerPath && (!is_pow2(lanesPerPath) || lanesPerPath > 64))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_schedule: lanesPerPath must be a power of two <= 64");
    if (blockSize && (blockSize % 64 != 0 || blockSize > 1024))
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_schedule: blockSize must be a multiple of 64 <= 1024");
    if (blocksPerCU > 32) return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_schedule: blocksPerCU <= 32");
    ctx->adjG = lanesPerPath;
    ctx->adjBlock = blockSize;
    ctx->adjBpc = blocksPerCU;
    return TWISTY_HIP_OK;
}

int twisty_hip_run_adjoint(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset,
      uint32_t perturbationsPerPath, uint64_t seed, twisty_hip_partials *out, double *dLog10EstimateDTable,
      double *dLog10EstimateDSegments)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!out || !dLog10EstimateDTable) return fail(ctx, TWISTY_HIP_ERR_ARG, "run_adjoint: out and dTable are required");
    const bool sweep = dLog10EstimateDSegments != nullptr;
    // The roulette applies to the exit-direction sweep only; the table
    // gradient then comes from a table-only pass over every path.
    const bool roulette = sweep && ctx->adjKappa > 0.0f;
    FwdRequest r;
    r.numPaths = numPaths;
    r.pathOffset = pathOffset;
    r.seed = seed;
    r.K = perturbationsPerPath;
    r.chainLength = 1;
    r.pathLog2 = true;
    r.finalSeg = true;  // every adjoint pass starts from the paths' final segments
    const uint32_t N = ctx->N;
    if (roulette && numPaths > 0xFFFFFFFFull)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "run_adjoint: the roulette handles at most 2^32 - 1 paths");
    // Schedules: the user's adjoint schedule applies to the pass that has the
    // sweep (or to the table pass when there is no sweep); the table pass in
    // front of a roulette sweep takes the built-in choice.
    uint32_t G = ctx->adjG, B = ctx->adjBlock, bpc = ctx->adjBpc, lds = 0;
    uint32_t tG = 0, tB = 0, tBpc = 0, tLds = 0;
    if (ctx->haveCurve && ctx->haveTables) {
        std::string why;
        int rc0 = resolve_adjoint_schedule(ctx, N, sweep, G, B, bpc, lds, why);
        if (!rc0 && roulette) rc0 = resolve_adjoint_schedule(ctx, N, false, tG, tB, tBpc, tLds, why);
        if (rc0) return fail(ctx, rc0, "run_adjoint: " + why);
    }
    const bool dirs = sweep && ctx->numDirs > 0 && ctx->dirsN == N;
    int rc = enqueue_forward(ctx, r, "run_adjoint");
    if (rc) return rc;
    const uint32_t grid = ctx->numCUs * bpc;
    const uint32_t tGrid = roulette ? ctx->numCUs * tBpc : grid;
    const uint32_t nb = ctx->tableSize;
    if ((rc = ensure_dev(ctx, ctx->dGradTPartial, ctx->gradTPartialCap, (uint64_t)tGrid * nb, "table gradient partials")))
        return rc;
    if ((rc = ensure_dev(ctx, ctx->dGradT, ctx->gradTCap, nb, "table gradient"))) return rc;
    if (sweep) {
        if ((rc = ensure_dev(ctx, ctx->dGradSPartial, ctx->gradSPartialCap, (uint64_t)grid * 3ull * N,
                   "segment gradient partials")))
            return rc;
        if ((rc = ensure_dev(ctx, ctx->dGradS, ctx->gradSCap, 3ull * N, "segment gradient"))) return rc;
    }
    if (dirs && (rc = ensure_dev(ctx, ctx->dDirStats, ctx->dirStatsCap, (uint64_t)grid * kDirStatsMax,
                       "direction statistics")))
        return rc;
    AdjParams a {};
    a.finalSeg = ctx->dFinalSeg;
    a.pathLog2 = ctx->dPathLog2;
    a.fin = ctx->dFinal;
    a.table = ctx->dTable;
    a.gradTablePartial = ctx->dGradTPartial;
    a.seed = seed;
    a.pathOffset = pathOffset;
    a.numPaths = numPaths;
    a.N = N;
    a.K = perturbationsPerPath;
    a.maxDiff = (N - 2 < 25u) ? (N - 2) : 25u;
    a.stream = 0;
    a.tableSize = (int32_t)nb;
    a.tableStride = table_stride(nb);
    a.segStride = (3u * N + 63u) & ~63u;
    a.minK = ctx->minK;
    a.maxK = ctx->maxK;
    a.step = ctx->step;
    a.invDs = (float)N / ctx->arclength;
    a.invStep = 1.0f / ctx->step;
    hipError_t e;
    if (roulette) {
        // exact table gradient over every path
        AdjParams t = a;
        t.S = (N + tG - 1) / tG;
        t.scratchRow = forward_scratch_row(tG, N);
        t.tableGrad = 1;
        e = launch_adjoint(t, tG, false, tGrid, tB, tLds, ctx->stream);
        if (e != hipSuccess) return hipfail(ctx, e, "adjoint table kernel launch");
        if ((rc = ensure_dev(ctx, ctx->dPick, ctx->pickCap, numPaths, "roulette list"))) return rc;
        if ((rc = ensure_dev(ctx, ctx->dPickCount, ctx->pickCountCap, 1, "roulette count"))) return rc;
        HIP_TRY(ctx, hipMemsetAsync(ctx->dPickCount, 0, sizeof(uint32_t), ctx->stream));
        e = launch_adjoint_pick(ctx->dPathLog2, ctx->dFinal, numPaths, pathOffset, seed, ctx->adjKappa, ctx->dPick,
              ctx->dPickCount, ctx->stream);
        if (e != hipSuccess) return hipfail(ctx, e, "roulette kernel launch");
    }
    a.pick = roulette ? ctx->dPick : nullptr;
    a.pickCount = roulette ? ctx->dPickCount : nullptr;
    a.gradSegPartial = sweep ? ctx->dGradSPartial : nullptr;
    a.S = (N + G - 1) / G;
    a.scratchRow = forward_scratch_row(G, N);
    a.tableGrad = roulette ? 0u : 1u;
    a.dirs = dirs ? ctx->dDirs : nullptr;
    a.numDirs = dirs ? ctx->numDirs : 0u;
    a.dirStats = dirs ? ctx->dDirStats : nullptr;
    e = launch_adjoint(a, G, sweep, grid, B, lds, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "adjoint kernel launch");
    e = launch_grad_reduce(ctx->dGradTPartial, tGrid, nb, ctx->dFinal, ctx->dGradT, ctx->stream);
    if (e != hipSuccess) return hipfail(ctx, e, "gradient reduce launch");
    if (sweep) {
        e = launch_grad_reduce(ctx->dGradSPartial, grid, 3u * N, ctx->dFinal, ctx->dGradS, ctx->stream);
        if (e != hipSuccess) return hipfail(ctx, e, "gradient reduce launch");
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev3, ctx->stream));
    FinalOut fo;
    HIP_TRY(ctx, hipMemcpyAsync(&fo, ctx->dFinal, sizeof(FinalOut), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dLog10EstimateDTable, ctx->dGradT, nb * sizeof(double), hipMemcpyDeviceToHost,
                       ctx->stream));
    if (sweep)
        HIP_TRY(ctx, hipMemcpyAsync(dLog10EstimateDSegments, ctx->dGradS, 3ull * N * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    std::vector<float> stats(dirs ? (size_t)grid * kDirStatsMax : 0);
    if (dirs)
        HIP_TRY(ctx, hipMemcpyAsync(stats.data(), ctx->dDirStats, stats.size() * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    uint32_t picked = 0;
    if (roulette)
        HIP_TRY(ctx, hipMemcpyAsync(&picked, ctx->dPickCount, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->lastPicked = roulette ? picked : numPaths;
    ctx->lastDirs = 0;
    if (dirs) {
        // FP64 sums over blocks; E = log10(sum w / P) with w scaled by 2^-M:
        // dE/dv = G' / ln 10, G' = sum_p w_p dln w_p/dv / S.
        double sum[kDirStatsMax] = {};
        for (uint32_t blk = 0; blk < grid; ++blk)
            for (uint32_t k = 0; k < kDirStatsMax; ++k) sum[k] += (double)stats[(size_t)blk * kDirStatsMax + k];
        const double S = fo.sumScaled, ln10 = 2.30258509299404568402;
        for (uint32_t d = 0; d < ctx->numDirs; ++d) {
            double g = 0.0;
            for (uint64_t q = 0; q < 3ull * N; ++q)
                g += dLog10EstimateDSegments[q] * (double)ctx->dirsHost[(size_t)d * 3 * N + q];
            const double Gp = g * ln10;  // in ln units
            const double varR = sum[3 * d];  // a sum of x^2 (1 - q) >= 0: 0 for the exact sweep
            const double varM = std::max(0.0, sum[3 * d + 1] - 2.0 * Gp * sum[3 * d + 2] + Gp * Gp * sum[3 * kMaxDirs]);
            ctx->lastDirGrad[d] = g;
            ctx->lastDirRouletteStd[d] = (S > 0.0) ? std::sqrt(varR) / (S * ln10) : 0.0;
            ctx->lastDirMcStd[d] = (S > 0.0) ? std::sqrt(varM) / (S * ln10) : 0.0;
        }
        ctx->lastDirs = ctx->numDirs;
    }
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastFwdMs, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastTotalMs, ctx->ev0, ctx->ev3));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->lastAdjMs, ctx->ev2, ctx->ev3));
    ctx->lastHistPaths = 0;
    ctx->lastHistBins = 0;
    ctx->lastFinalPaths = numPaths;  // the forward pass above kept the final state
    copy_partials(fo, numPaths, out);
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

int twisty_hip_set_adjoint_directions(twisty_hip_ctx *ctx, const float *directions, uint32_t numDirections)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (numDirections == 0) {
        ctx->numDirs = 0;
        return TWISTY_HIP_OK;
    }
    if (!ctx->haveCurve) return fail(ctx, TWISTY_HIP_ERR_STATE, "set_adjoint_directions: set_curve first");
    if (!directions || numDirections > kMaxDirs)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_directions: need 1.." + std::to_string(kMaxDirs) +
                    " directions of 3N floats");
    const uint64_t n = (uint64_t)numDirections * 3ull * ctx->N;
    for (uint64_t q = 0; q < n; ++q)
        if (!std::isfinite(directions[q]))
            return fail(ctx, TWISTY_HIP_ERR_ARG, "set_adjoint_directions: non-finite entry");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_dev(ctx, ctx->dDirs, ctx->dirsCap, n, "adjoint directions"))) return rc;
    HIP_TRY(ctx, hipMemcpy(ctx->dDirs, directions, n * sizeof(float), hipMemcpyHostToDevice));
    ctx->dirsHost.assign(directions, directions + n);
    ctx->numDirs = numDirections;
    ctx->dirsN = ctx->N;
    return TWISTY_HIP_OK;
}

int twisty_hip_last_adjoint_direction_stats(twisty_hip_ctx *ctx, double *gradient, double *rouletteStd,
      double *monteCarloStd)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (ctx->lastDirs == 0)
        return fail(ctx, TWISTY_HIP_ERR_STATE, "last_adjoint_direction_stats: the last adjoint had no directions "
                    "(set_adjoint_directions, then run_adjoint with dSegments)");
    for (uint32_t d = 0; d < ctx->lastDirs; ++d) {
        if (gradient) gradient[d] = ctx->lastDirGrad[d];
        if (rouletteStd) rouletteStd[d] = ctx->lastDirRouletteStd[d];
        if (monteCarloStd) monteCarloStd[d] = ctx->lastDirMcStd[d];
    }
    return TWISTY_HIP_OK;
}

int twisty_hip_last_timing(twisty_hip_ctx *ctx, float *forwardMs, float *totalMs)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (forwardMs) *forwardMs = ctx->lastFwdMs;
    if (totalMs) *totalMs = ctx->lastTotalMs;
    return TWISTY_HIP_OK;
}

int twisty_hip_last_adjoint_timing(twisty_hip_ctx *ctx, float *adjointMs)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (adjointMs) *adjointMs = ctx->lastAdjMs;
    return TWISTY_HIP_OK;
}

}  // extern "C"
