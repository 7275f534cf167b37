// batch_capi.hip -- C ABI of the batched experiments (declared in
// include/twisty_hip.h): upload, schedule, work-item list, HIP-graph capture,
// replay and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/twisty_hip.h"
#include "batched.hpp"
#include "ctx.hpp"

struct twisty_hip_batch {
    twisty_hip_ctx *ctx = nullptr;
    std::vector<twisty_hip_experiment> exps;
    std::vector<BatchTable> tables;
    uint32_t numBins = 0;
    uint64_t totalPaths = 0;

    float *dTableData = nullptr;
    float *dSegs = nullptr;
    BatchTable *dTables = nullptr;
    BatchExp *dExps = nullptr;
    BatchItem *dItems = nullptr;
    BlockPartial *dItemPartials = nullptr;
    uint32_t *dQueue = nullptr;
    uint32_t *dExpItemStart = nullptr;
    uint32_t *dExpItems = nullptr;
    double *dAbsorb = nullptr;
    FinalOut *dOuts = nullptr;
    float *dPerPath = nullptr;
    uint64_t itemsCap = 0, perPathCap = 0;
    std::vector<uint32_t> segOffset;  // float offset of each experiment's segments
    std::vector<double> absorbLog2;   // per path: (N-1) log2 expf(-absorption ds)

    twisty_hip_batch_schedule sched {};
    uint32_t grid = 0, lds = 0, rows = 0, numItems = 0;
    bool planned = false;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graphPerPath = false;
};

namespace {

const twisty_hip_batch_schedule kDefaultSchedule { 256, 0, 16, 4096, 0 };
// Lanes of the batched kernel resident per CU by registers: three waves of 64
// per SIMD at 168 VGPRs (lean 0 and 2), two for the lean form (about 190).
constexpr uint32_t kBatchedLanesPerCU = 768;
constexpr uint32_t kBatchedLeanLanesPerCU = 512;

uint32_t lanes_for(uint32_t N, uint32_t spl)
{
    uint32_t G = 2;
    while (G < 64 && (N + G - 1) / G > spl) G <<= 1;
    return G;
}

void drop_graph(twisty_hip_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Builds experiment descriptors, work items (longest first) and the
// per-experiment item lists for the current schedule, and uploads them.
int plan(twisty_hip_batch *b)
{
    twisty_hip_ctx *ctx = b->ctx;
    const twisty_hip_batch_schedule &sc = b->sched;
    if (sc.segments_per_lane < 1 || sc.segments_per_lane > 64 || sc.tile_paths < 1 || sc.lean > 2)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "batch: segmentsPerLane in [1, 64], tilePaths >= 1, lean 0, 1 or 2");
    if (!batched_block_supported(sc.block_size, (int)sc.lean))
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED,
              "batch: blockSize must be 256 or 512 (lean 0 or 1), 256 (lean 2)");
    const uint32_t nE = (uint32_t)b->exps.size();
    std::vector<BatchExp> de(nE);
    uint32_t rows = 0;
    uint64_t pathOut = 0;
    for (uint32_t e = 0; e < nE; ++e) {
        const twisty_hip_experiment &x = b->exps[e];
        const uint32_t N = x.num_segments;
        const uint32_t G = lanes_for(N, sc.segments_per_lane);
        if ((N + G - 1) / G > sc.segments_per_lane)
            return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED,
                  "batch: experiment " + std::to_string(e) + " has N=" + std::to_string(N) +
                        " > 64 * segmentsPerLane");
        BatchExp &d = de[e];
        d.numPaths = x.num_paths;
        d.seed = x.seed;
        d.pathOut = pathOut;
        d.N = N;
        d.K = x.perturbations_per_path;
        d.maxDiff = (N - 2 < 25u) ? (N - 2) : 25u;
        d.G = G;
        d.S = (N + G - 1) / G;
        d.scratchRow = forward_scratch_row(G, N);
        d.segOffset = b->segOffset[e];
        d.table = x.table_index;
        d.invDs = (float)N / x.arclength;  // CurvePerturbUtils.cpp:223
        d.absorbLog10PerPath = (float)(b->absorbLog2[e] * 0.30102999566398119521);
        d.chainLength = std::max<uint32_t>(1u, x.chain_length);
        d.burnIn = x.burn_in;
        rows = std::max(rows, d.scratchRow + 1);
        pathOut += x.num_paths;
    }
    const uint32_t ts = table_stride(b->numBins);
    const uint32_t lds = batched_lds_bytes(sc.block_size, rows, ts, (int)sc.lean);
    if (lds > ctx->ldsPerBlock)
        return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED,
              "batch: schedule needs " + std::to_string(lds) + " LDS bytes per block, limit " +
                    std::to_string(ctx->ldsPerBlock));
    uint32_t bpc = sc.blocks_per_cu;
    if (bpc == 0) {
        bpc = ctx->ldsPerCU / lds;
        // registers: three waves per SIMD (168 VGPRs), two for lean (it is
        // not capped: it would spill)
        const uint32_t lanes = sc.lean == 1 ? kBatchedLeanLanesPerCU : kBatchedLanesPerCU;
        if (bpc * sc.block_size > lanes) bpc = lanes / sc.block_size;
        if (bpc < 1) bpc = 1;
    }
    // work items of whole chains (about tilePaths paths each), longest first
    // (cost ~ paths * (K * sub-chain length + N) plus the burn-in)
    std::vector<BatchItem> items;
    for (uint32_t e = 0; e < nE; ++e) {
        const uint64_t L = de[e].chainLength;
        const uint64_t chains = (de[e].numPaths + L - 1) / L;
        const uint64_t per = std::max<uint64_t>(1, sc.tile_paths / L);
        for (uint64_t c0 = 0; c0 < chains; c0 += per) {
            BatchItem it;
            it.begin = c0;
            it.exp = e;
            it.count = (uint32_t)std::min<uint64_t>(per, chains - c0);
            items.push_back(it);
        }
    }
    auto cost = [&](const BatchItem &it) {
        const BatchExp &d = de[it.exp];
        const double perturb = (double)d.K * (double)(d.maxDiff + 2) * 0.5 * (double)d.G;
        const double paths = std::min<double>((double)it.count * d.chainLength, (double)d.numPaths);
        return paths * (perturb + (double)d.N) + (double)it.count * d.burnIn * perturb;
    };
    std::stable_sort(items.begin(), items.end(), [&](const BatchItem &a, const BatchItem &c) { return cost(a) > cost(c); });
    std::vector<uint32_t> start(nE + 1, 0), list(items.size());
    for (const BatchItem &it : items) ++start[it.exp + 1];
    for (uint32_t e = 0; e < nE; ++e) start[e + 1] += start[e];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t i = 0; i < items.size(); ++i) list[fill[items[i].exp]++] = i;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_dev(ctx, b->dItems, b->itemsCap, items.size(), "batch items"))) return rc;
    uint64_t cap2 = 0;
    (void)hipFree(b->dItemPartials);
    b->dItemPartials = nullptr;
    if ((rc = ensure_dev(ctx, b->dItemPartials, cap2, items.size(), "batch item partials"))) return rc;
    uint64_t cap3 = 0;
    (void)hipFree(b->dExpItems);
    b->dExpItems = nullptr;
    if ((rc = ensure_dev(ctx, b->dExpItems, cap3, items.size(), "batch item lists"))) return rc;
    HIP_TRY(ctx, hipMemcpy(b->dItems, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(b->dExpItems, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(b->dExpItemStart, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(b->dExps, de.data(), de.size() * sizeof(BatchExp), hipMemcpyHostToDevice));
    HIP_TRY(ctx, batched_prepare(sc.block_size, (int)sc.lean, lds));
    b->grid = ctx->numCUs * bpc;
    b->lds = lds;
    b->rows = rows;
    b->numItems = (uint32_t)items.size();
    b->planned = true;
    drop_graph(b);
    return TWISTY_HIP_OK;
}

int capture(twisty_hip_batch *b, bool perPath)
{
    twisty_hip_ctx *ctx = b->ctx;
    drop_graph(b);
    int rc;
    if (perPath && (rc = ensure_dev(ctx, b->dPerPath, b->perPathCap, b->totalPaths, "batch per-path buffer")))
        return rc;
    BatchParams p {};
    p.exps = b->dExps;
    p.items = b->dItems;
    p.tables = b->dTables;
    p.tableData = b->dTableData;
    p.segs = b->dSegs;
    p.itemPartials = b->dItemPartials;
    p.perPathLog10 = perPath ? b->dPerPath : nullptr;
    p.queue = b->dQueue;
    p.numItems = b->numItems;
    p.rows = b->rows;
    p.tableStride = table_stride(b->numBins);
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch_batched(p, b->grid, b->sched.block_size, (int)b->sched.lean, b->lds,
          b->dExpItemStart, b->dExpItems, b->dAbsorb, (uint32_t)b->exps.size(), b->dOuts, ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hipfail(ctx, le, "batched launch during capture");
    }
    if (ce != hipSuccess) return hipfail(ctx, ce, "hipStreamEndCapture");
    b->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graphPerPath = perPath;
    return TWISTY_HIP_OK;
}

int replay(twisty_hip_batch *b, float *ms)
{
    twisty_hip_ctx *ctx = b->ctx;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, hipGraphLaunch(b->exec, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    return TWISTY_HIP_OK;
}

}  // namespace

namespace twisty_hip {

void batch_release(twisty_hip_batch *b)
{
    if (!b || !b->ctx) return;
    twisty_hip_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph(b);
    for (void *p : { (void *)b->dTableData, (void *)b->dSegs, (void *)b->dTables, (void *)b->dExps,
               (void *)b->dItems, (void *)b->dItemPartials, (void *)b->dQueue, (void *)b->dExpItemStart,
               (void *)b->dExpItems, (void *)b->dAbsorb, (void *)b->dOuts, (void *)b->dPerPath })
        (void)hipFree(p);
    b->dTableData = b->dSegs = b->dPerPath = nullptr;
    b->dTables = nullptr;
    b->dExps = nullptr;
    b->dItems = nullptr;
    b->dItemPartials = nullptr;
    b->dQueue = b->dExpItemStart = b->dExpItems = nullptr;
    b->dAbsorb = nullptr;
    b->dOuts = nullptr;
    auto &v = ctx->batches;
    v.erase(std::remove(v.begin(), v.end(), b), v.end());
    b->ctx = nullptr;
}

}  // namespace twisty_hip

extern "C" {

int twisty_hip_batch_create(twisty_hip_ctx *ctx, const twisty_hip_experiment *experiments, uint32_t numExperiments,
      const float *tables, uint32_t numBins, uint32_t numTables, const float *tableMeta, const float *positions,
      uint64_t numPoints, twisty_hip_batch **out)
{
    if (!ctx) return TWISTY_HIP_ERR_ARG;
    if (!out) return fail(ctx, TWISTY_HIP_ERR_ARG, "batch_create: out is NULL");
    *out = nullptr;
    if (!experiments || numExperiments == 0 || !tables || numBins < 2 || numBins > 65536 || numTables == 0 ||
          !tableMeta || !positions)
        return fail(ctx, TWISTY_HIP_ERR_ARG, "batch_create: need experiments, tables (2..65536 bins), meta, positions");
    auto *b = new twisty_hip_batch();
    b->ctx = ctx;
    ctx->batches.push_back(b);
    b->numBins = numBins;
    b->exps.assign(experiments, experiments + numExperiments);
    auto bail = [&](int code, const std::string &m) {
        twisty_hip_batch_destroy(b);
        return fail(ctx, code, "batch_create: " + m);
    };
    // tables
    std::vector<float> absorptionOf(numTables);
    for (uint32_t t = 0; t < numTables; ++t) {
        const float *m = tableMeta + 5 * t;
        if (!(m[2] > 0.0f) || !std::isfinite(m[0]) || !std::isfinite(m[1]) || !(m[1] >= m[0]) ||
              !std::isfinite(m[3]) || !std::isfinite(m[4]))
            return bail(TWISTY_HIP_ERR_ARG, "bad grid of table " + std::to_string(t));
        for (uint32_t i = 0; i < numBins; ++i)
            if (std::isinf(tables[(uint64_t)t * numBins + i])) return bail(TWISTY_HIP_ERR_ARG, "table contains inf");
        BatchTable bt;
        bt.offset = t * table_stride(numBins);
        bt.size = (int32_t)numBins;
        bt.minK = m[0];
        bt.maxK = m[1];
        bt.step = m[2];
        b->tables.push_back(bt);
        absorptionOf[t] = m[4];
    }
    // experiments and their initial segments (FP64 differences rounded once)
    std::vector<float> segs;
    std::vector<double> absorb(numExperiments);
    for (uint32_t e = 0; e < numExperiments; ++e) {
        const twisty_hip_experiment &x = experiments[e];
        const uint32_t N = x.num_segments;
        if (N < 4 || x.num_paths == 0 || x.perturbations_per_path == 0 || x.table_index >= numTables ||
              !(x.arclength > 0.0f) || (uint64_t)x.positions_offset + N >= numPoints ||
              x.num_paths > (1ull << 40) || x.burn_in > (1u << 24))
            return bail(TWISTY_HIP_ERR_ARG, "bad experiment " + std::to_string(e));
        b->segOffset.push_back((uint32_t)segs.size());
        for (uint32_t i = 0; i < N; ++i) {
            for (int a = 0; a < 3; ++a) {
                const uint64_t q = (uint64_t)(x.positions_offset + i) * 3 + a;
                segs.push_back((float)((double)positions[q + 3] - (double)positions[q]));
            }
        }
        // PathWeighters.cpp:37: float argument, float exp, per joint, with the
        // experiment's own ds (computed in float, Curve.cpp:24) and the table's
        // absorption coefficient
        const float dsE = x.arclength / (float)N;
        absorb[e] = (double)(N - 1) * std::log2((double)std::exp((float)(-absorptionOf[x.table_index] * dsE)));
        b->totalPaths += x.num_paths;
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const uint64_t ts = table_stride(numBins);
    std::vector<float> tdata(ts * numTables, 0.0f);
    for (uint32_t t = 0; t < numTables; ++t)
        std::copy(tables + (uint64_t)t * numBins, tables + (uint64_t)(t + 1) * numBins, tdata.begin() + t * ts);
    hipError_t he;
    if ((he = hipMalloc(&b->dTableData, tdata.size() * sizeof(float))) != hipSuccess ||
          (he = hipMalloc(&b->dSegs, segs.size() * sizeof(float))) != hipSuccess ||
          (he = hipMalloc(&b->dTables, b->tables.size() * sizeof(BatchTable))) != hipSuccess ||
          (he = hipMalloc(&b->dExps, numExperiments * sizeof(BatchExp))) != hipSuccess ||
          (he = hipMalloc(&b->dQueue, sizeof(uint32_t))) != hipSuccess ||
          (he = hipMalloc(&b->dExpItemStart, (numExperiments + 1) * sizeof(uint32_t))) != hipSuccess ||
          (he = hipMalloc(&b->dAbsorb, numExperiments * sizeof(double))) != hipSuccess ||
          (he = hipMalloc(&b->dOuts, numExperiments * sizeof(FinalOut))) != hipSuccess) {
        const int code = hipfail(ctx, he, "batch_create: hipMalloc");
        twisty_hip_batch_destroy(b);
        return code;
    }
    if ((he = hipMemcpy(b->dTableData, tdata.data(), tdata.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
          (he = hipMemcpy(b->dSegs, segs.data(), segs.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess ||
          (he = hipMemcpy(b->dTables, b->tables.data(), b->tables.size() * sizeof(BatchTable), hipMemcpyHostToDevice)) !=
                hipSuccess ||
          (he = hipMemcpy(b->dAbsorb, absorb.data(), absorb.size() * sizeof(double), hipMemcpyHostToDevice)) !=
                hipSuccess) {
        const int code = hipfail(ctx, he, "batch_create: upload");
        twisty_hip_batch_destroy(b);
        return code;
    }
    b->absorbLog2 = absorb;
    b->sched = kDefaultSchedule;
    const int rc = plan(b);
    if (rc) {
        twisty_hip_batch_destroy(b);
        return rc;
    }
    *out = b;
    return TWISTY_HIP_OK;
}

int twisty_hip_batch_destroy(twisty_hip_batch *b)
{
    if (!b) return TWISTY_HIP_ERR_ARG;
    batch_release(b);
    delete b;
    return TWISTY_HIP_OK;
}

int twisty_hip_batch_set_schedule(twisty_hip_batch *b, const twisty_hip_batch_schedule *s)
{
    if (!b || !s) return TWISTY_HIP_ERR_ARG;
    if (!b->ctx) return TWISTY_HIP_ERR_STATE;  // its context was destroyed
    const twisty_hip_batch_schedule old = b->sched;
    b->sched = *s;
    if (b->sched.block_size == 0) b->sched.block_size = kDefaultSchedule.block_size;
    if (b->sched.segments_per_lane == 0) b->sched.segments_per_lane = kDefaultSchedule.segments_per_lane;
    if (b->sched.tile_paths == 0) b->sched.tile_paths = kDefaultSchedule.tile_paths;
    const int rc = plan(b);
    if (rc) {
        b->sched = old;
        const std::string msg = b->ctx->lastError;
        (void)plan(b);
        b->ctx->lastError = msg;
    }
    return rc;
}

int twisty_hip_batch_get_schedule(twisty_hip_batch *b, twisty_hip_batch_schedule *s)
{
    if (!b || !s) return TWISTY_HIP_ERR_ARG;
    if (!b->ctx) return TWISTY_HIP_ERR_STATE;  // its context was destroyed
    *s = b->sched;
    if (s->blocks_per_cu == 0 && b->ctx->numCUs) s->blocks_per_cu = b->grid / b->ctx->numCUs;
    return TWISTY_HIP_OK;
}

int twisty_hip_batch_run(twisty_hip_batch *b, twisty_hip_partials *out, float *perPathLog10, float *deviceMs)
{
    if (!b) return TWISTY_HIP_ERR_ARG;
    if (!b->ctx) return TWISTY_HIP_ERR_STATE;  // its context was destroyed
    twisty_hip_ctx *ctx = b->ctx;
    if (!out) return fail(ctx, TWISTY_HIP_ERR_ARG, "batch_run: out is NULL");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    const bool perPath = perPathLog10 != nullptr;
    if (!b->exec || b->graphPerPath != perPath)
        if ((rc = capture(b, perPath))) return rc;
    float ms = 0.0f;
    if ((rc = replay(b, &ms))) return rc;
    const uint32_t nE = (uint32_t)b->exps.size();
    std::vector<FinalOut> fo(nE);
    HIP_TRY(ctx, hipMemcpy(fo.data(), b->dOuts, nE * sizeof(FinalOut), hipMemcpyDeviceToHost));
    if (perPath)
        HIP_TRY(ctx, hipMemcpy(perPathLog10, b->dPerPath, b->totalPaths * sizeof(float), hipMemcpyDeviceToHost));
    for (uint32_t e = 0; e < nE; ++e) {
        twisty_hip_partials &o = out[e];
        o.log10_sum_w = fo[e].log10SumW;
        o.log10_sum_w2 = fo[e].log10SumW2;
        o.max_log10_w = fo[e].maxLog10W;
        o.num_paths = b->exps[e].num_paths;
        o.num_valid = fo[e].nValid;
        o.num_invalid = fo[e].nInvalid;
        o.num_clamped = fo[e].nClamped;
    }
    if (deviceMs) *deviceMs = ms;
    ctx->lastFwdMs = ms;
    ctx->lastTotalMs = ms;
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

int twisty_hip_batch_autotune(twisty_hip_batch *b, uint32_t reps, twisty_hip_batch_schedule *best,
      twisty_hip_batch_schedule *tried, float *triedMs, uint32_t cap, uint32_t *numTried)
{
    if (!b) return TWISTY_HIP_ERR_ARG;
    if (!b->ctx) return TWISTY_HIP_ERR_STATE;  // its context was destroyed
    twisty_hip_ctx *ctx = b->ctx;
    if (reps == 0) reps = 1;
    // Candidates: block size x lanes policy x blocks per CU x tile size with
    // the tables in LDS; the tables in global memory at block 256 with as many
    // blocks per CU as fit, where 15 segments per lane keep every state within
    // 16 rows, so three blocks share a CU (the C3 batch: 127 ms against 149 ms
    // for the best table-in-LDS schedule), and the lean form.  A candidate
    // whose LDS does not fit is skipped.
    std::vector<twisty_hip_batch_schedule> cand;
    for (uint32_t block : { 256u, 512u })
        for (uint32_t spl : { 8u, 16u })
            for (uint32_t bpc : { 1u, 2u })
                for (uint32_t tile : { 1024u, 4096u, 16384u })
                    cand.push_back({ block, bpc, spl, tile, 0u });
    for (uint32_t spl : { 8u, 12u, 15u, 16u })
        for (uint32_t tile : { 1024u, 4096u, 16384u })
            cand.push_back({ 256u, 0u, spl, tile, 2u });
    for (uint32_t spl : { 8u, 16u })
        cand.push_back({ 256u, 0u, spl, 4096u, 1u });
    twisty_hip_batch_schedule bestS = b->sched;
    float bestMs = INFINITY;
    uint32_t n = 0;
    for (const auto &c : cand) {
        b->sched = c;
        if (plan(b)) continue;
        if (c.blocks_per_cu * c.block_size > 2048 || (uint64_t)c.blocks_per_cu * b->lds > ctx->ldsPerCU) continue;
        if (capture(b, false)) continue;
        float m = INFINITY;
        bool ok = true;
        for (uint32_t r = 0; r <= reps; ++r) {  // the first replay warms up
            float t = 0.0f;
            if (replay(b, &t)) {
                ok = false;
                break;
            }
            if (r > 0) m = std::min(m, t);
        }
        if (!ok) continue;
        if (n < cap) {
            if (tried) tried[n] = c;
            if (triedMs) triedMs[n] = m;
        }
        ++n;
        if (m < bestMs) {
            bestMs = m;
            bestS = c;
        }
    }
    if (numTried) *numTried = n;
    if (!(bestMs < INFINITY)) return fail(ctx, TWISTY_HIP_ERR_UNSUPPORTED, "batch_autotune: no candidate ran");
    b->sched = bestS;
    int rc = plan(b);
    if (rc) return rc;
    if (best) *best = bestS;
    ctx->lastError.clear();
    return TWISTY_HIP_OK;
}

}  // extern "C"
