// batch_runner.cpp -- BatchedExperimentRunner: the experiments of a
// multi-experiment driver in one twisty_hip_batch (persistent kernel, HIP
// graph), each experiment's estimate as FullExperimentRunnerOptimalPerturb_HIP
// gives it (same table, initial curve, normaliser, draws, chains, burn-in).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <ctime>
#include <map>

#include "../../../include/twisty_hip.h"
#include "twisty_host.hpp"

namespace twisty {

BatchedExperimentRunner::BatchedExperimentRunner(const ExperimentRunner::ExperimentParameters &experimentParams)
    : m_ep(experimentParams)
{
}

std::optional<std::vector<ExperimentRunner::ExperimentResults>> BatchedExperimentRunner::RunExperiments(
      const std::vector<Experiment> &experiments)
{
    m_lastError.clear();
    auto fail = [&](const std::string &msg) -> std::optional<std::vector<ExperimentRunner::ExperimentResults>> {
        m_lastError = msg;
        return std::nullopt;
    };
    const ExperimentRunner::ExperimentParameters &ep = m_ep;
    if (experiments.empty()) return fail("no experiments");
    if (ep.weightingParameters.weightingMethod != WeightingMethod::RadiativeTransfer)
        return fail("only the radiative-transfer weighting (weightFunction 0) runs on the HIP kernels");
    if (ep.perturbMethod != ExperimentRunner::PerturbMethod::GeometricRandom) return fail("Invalid Perturb Method");
    if (ep.numPathsToSkip > (1u << 24)) return fail("pathsToSkip above 2^24");
    const uint32_t N = ep.numSegmentsPerCurve;
    if (N < 4) return fail("numSegments must be >= 4");
    const uint64_t nPaths = ep.numPathsInExperiment;
    if (nPaths == 0) return fail("pathsToGenerate must be >= 1");
    const auto t0 = std::chrono::high_resolution_clock::now();

    // One table per distinct ds (float bits), built once.
    std::map<uint32_t, uint32_t> tableOf;
    std::vector<float> tables, meta;
    std::vector<twisty_hip_experiment> exps(experiments.size());
    std::vector<float> positions;
    std::vector<double> normLog10(experiments.size());
    uint32_t numBins = 0;
    // Chain length 0: about 65536 chains over the whole batch.
    uint64_t L = ep.chainLength;
    if (L == 0) L = std::max<uint64_t>(1, (nPaths * experiments.size() + 65535) / 65536);
    if (L > 0xFFFFFFFFull) return fail("chain length above 2^32 - 1");
    for (size_t e = 0; e < experiments.size(); ++e) {
        const PerturbUtils::BoundaryConditions &bc = experiments[e].boundaryConditions;
        if (!(bc.arclength > 0.0f)) return fail("experiment " + std::to_string(e) + ": arclength must be positive");
        const float ds = bc.arclength / (float)N;
        uint32_t bits;
        std::memcpy(&bits, &ds, sizeof(bits));
        auto it = tableOf.find(bits);
        if (it == tableOf.end()) {
            PathWeighting::WeightLookupTableIntegral table(ep.weightingParameters, ds);
            try {
                table.Initialize(ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0);
            } catch (const std::exception &ex) {
                return fail(ex.what());
            }
            const auto &tv = table.AccessLookupTable();
            numBins = (uint32_t)tv.size();
            tables.insert(tables.end(), tv.begin(), tv.end());
            meta.insert(meta.end(), { table.GetMinCurvature(), table.GetMaxCurvature(), table.GetCurvatureStepSize(),
                                          table.GetDs(), ep.weightingParameters.absorption });
            it = tableOf.emplace(bits, (uint32_t)tableOf.size()).first;
        }
        std::vector<Vec3> arc;
        if (!PathGeneration::InitialArc(bc, N, arc))
            return fail("experiment " + std::to_string(e) + ": no initial curve: arclength " +
                        std::to_string(bc.arclength) + " is below the minimum " +
                        std::to_string(PathGeneration::CalculateMinimumArclength(bc, N)));
        normLog10[e] = PathWeighting::NormLog10((int)N, ds, bc);
        twisty_hip_experiment &x = exps[e];
        x.num_paths = nPaths;
        x.seed = experiments[e].seed ? experiments[e].seed : (uint64_t)std::time(nullptr);
        x.num_segments = N;
        x.perturbations_per_path = ep.perturbationsPerPath;
        x.table_index = it->second;
        x.positions_offset = (uint32_t)(positions.size() / 3);
        x.arclength = bc.arclength;
        x.chain_length = (uint32_t)L;
        x.burn_in = (uint32_t)ep.numPathsToSkip;
        for (const Vec3 &v : arc) positions.insert(positions.end(), { v.x, v.y, v.z });
    }
    m_numTables = (uint32_t)tableOf.size();
    const auto t1 = std::chrono::high_resolution_clock::now();

    twisty_hip_ctx *ctx = nullptr;
    if (twisty_hip_create(ep.device, &ctx) != TWISTY_HIP_OK)
        return fail("twisty_hip_create failed (device " + std::to_string(ep.device) + ")");
    twisty_hip_batch *batch = nullptr;
    struct Guard {
        twisty_hip_ctx *c;
        twisty_hip_batch **b;
        ~Guard()
        {
            if (*b) twisty_hip_batch_destroy(*b);
            twisty_hip_destroy(c);
        }
    } guard { ctx, &batch };
    int rc = twisty_hip_batch_create(ctx, exps.data(), (uint32_t)exps.size(), tables.data(), numBins, m_numTables,
          meta.data(), positions.data(), positions.size() / 3, &batch);
    if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    if (ep.autotune) {
        rc = twisty_hip_batch_autotune(batch, 1, nullptr, nullptr, nullptr, 0, nullptr);
        if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    }
    std::vector<twisty_hip_partials> parts(exps.size());
    float ms = 0.0f;
    rc = twisty_hip_batch_run(batch, parts.data(), nullptr, &ms);
    if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    m_deviceMs = ms;
    const auto t2 = std::chrono::high_resolution_clock::now();

    std::vector<ExperimentRunner::ExperimentResults> out(exps.size());
    const auto setupMs = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    const auto runMs = std::chrono::duration_cast<std::chrono::milliseconds>(t2 - t1).count();
    for (size_t e = 0; e < exps.size(); ++e) {
        ExperimentRunner::ExperimentResults &r = out[e];
        r.experimentWeight.log10 = parts[e].log10_sum_w + normLog10[e];
        r.totalPathsGenerated = nPaths;
        r.numFailedPaths = parts[e].num_invalid;
        r.log10SumSquares = parts[e].log10_sum_w2 + 2.0 * normLog10[e];
        r.kernelMs = ms;
        r.numClamped = parts[e].num_clamped;
        r.setupExperimentMs = setupMs;
        r.perturbExperimentMs = runMs;
        r.totalExperimentMs = setupMs + runMs;
    }
    return out;
}

}  // namespace twisty
