// path_generation.cpp -- ExperimentBase::MSegmentPathGenerationMC
// (Experiments/ExperimentBase.cpp:1246-1386) on the HIP independent-path
// kernel: the table of ds = arclength / N, the boundary, one
// twisty_hip_run_generated call, and the reference's Result.
#include <algorithm>
#include <cmath>

#include "../../../include/twisty_hip.h"
#include "twisty_host.hpp"

namespace twisty {
namespace ExperimentBase {

std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &bc, const ExperimentRunner::ExperimentParameters &ep,
      double pathNormalizerLog10, std::string &error)
{
    error.clear();
    if (ep.weightingParameters.weightingMethod != WeightingMethod::RadiativeTransfer) {
        error = "only the radiative-transfer weighting (weightFunction 0) runs on the HIP kernels";
        return std::nullopt;
    }
    const uint32_t N = ep.numSegmentsPerCurve;
    if (!(bc.arclength > 0.0f) || N < 4 || numExperimentPaths == 0) {
        error = "need arclength > 0, numSegments >= 4 and at least one path";
        return std::nullopt;
    }
    const float ds = bc.arclength / (float)N;  // ExperimentBase.cpp:1253
    PathWeighting::WeightLookupTableIntegral table(ep.weightingParameters, ds);
    try {
        table.Initialize(ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0);
    } catch (const std::exception &e) {
        error = e.what();
        return std::nullopt;
    }
    // set_curve wants a curve: the straight chord when no arc exists (only its
    // boundary is read by the generated run)
    std::vector<Vec3> arc;
    if (!PathGeneration::InitialArc(bc, N, arc)) {
        arc.resize(N + 1);
        for (uint32_t i = 0; i <= N; ++i)
            arc[i] = bc.m_startPos + (bc.m_endPos - bc.m_startPos) * ((float)i / (float)N);
    }
    twisty_hip_ctx *ctx = nullptr;
    if (twisty_hip_create(ep.device, &ctx) != TWISTY_HIP_OK) {
        error = "twisty_hip_create failed (device " + std::to_string(ep.device) + ")";
        return std::nullopt;
    }
    struct Guard {
        twisty_hip_ctx *c;
        ~Guard() { twisty_hip_destroy(c); }
    } guard { ctx };
    const auto &tv = table.AccessLookupTable();
    int rc = twisty_hip_set_tables(ctx, tv.data(), (uint32_t)tv.size(), 1, table.GetMinCurvature(),
          table.GetMaxCurvature(), table.GetCurvatureStepSize(), table.GetDs(), ep.weightingParameters.absorption);
    const auto bc13 = PerturbUtils::ToCudaSafe(bc);
    if (rc == TWISTY_HIP_OK) rc = twisty_hip_set_curve(ctx, reinterpret_cast<const float *>(arc.data()), N, bc13.data());
    twisty_hip_partials part {};
    double mn = INFINITY;
    if (rc == TWISTY_HIP_OK) rc = twisty_hip_run_generated(ctx, numExperimentPaths, 0, seed, &part, &mn, nullptr, nullptr);
    if (rc != TWISTY_HIP_OK) {
        error = twisty_hip_last_error(ctx);
        return std::nullopt;
    }
    float fwdMs = 0.0f, totMs = 0.0f;
    twisty_hip_last_timing(ctx, &fwdMs, &totMs);
    Result r;
    r.numValidPaths = part.num_valid;
    r.numPathsTotal = numExperimentPaths;
    r.totalWeight.log10 = part.num_valid
          ? part.log10_sum_w + pathNormalizerLog10 - std::log10((double)part.num_valid)
          : -INFINITY;  // the reference returns 0 without valid paths (:1380-1384)
    r.minPathWeightLog10 = mn + pathNormalizerLog10;
    r.maxPathWeightLog10 = part.max_log10_w + pathNormalizerLog10;
    r.kernelMs = totMs;
    return r;
}

float QuadraticMinimumArclength(const PerturbUtils::BoundaryConditions &g, uint32_t M)
{
    const Vec3 ns = g.m_startDir + g.m_endDir;
    const Vec3 dx = g.m_endPos - g.m_startPos;
    const float Mf = (float)M;
    const float a = ns.Dot(ns) - ((Mf - 4.0f) * Mf + 4.0f);
    const float b = -2.0f * Mf * ns.Dot(dx);
    const float c = Mf * Mf * dx.Dot(dx);
    const float r1 = (-b - std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    const float r2 = (-b + std::sqrt(b * b - 4.0f * a * c)) / (2.0f * a);
    float best = 1000000.0f;
    if (!std::isnan(r1) && r1 > 0.0f) best = std::min(best, r1);
    if (!std::isnan(r2) && r2 > 0.0f) best = std::min(best, r2);
    return best;
}

std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &bc, const ExperimentRunner::ExperimentParameters &ep,
      double pathNormalizerLog10, const PathWeighting::CachedMultiArclengthWeightLookupTable &tables, float maxDs,
      std::string &error)
{
    error.clear();
    const uint32_t N = ep.numSegmentsPerCurve;
    if (N < 4 || numExperimentPaths == 0 || tables.Count() == 0) {
        error = "need numSegments >= 4, at least one path and one cached table";
        return std::nullopt;
    }
    const float minDs = QuadraticMinimumArclength(bc, N) / (float)N;
    if (maxDs < minDs) return Result {};  // "No solution for current environment"
    const size_t T = tables.Count();
    const uint32_t bins = (uint32_t)tables.Table(0).AccessLookupTable().size();
    std::vector<float> data, meta;
    data.reserve(T * bins);
    for (size_t t = 0; t < T; ++t) {
        const auto &tb = tables.Table(t);
        const auto &v = tb.AccessLookupTable();
        data.insert(data.end(), v.begin(), v.end());
        meta.insert(meta.end(), { tb.GetDs(), tb.GetMinCurvature(), tb.GetMaxCurvature(), tb.GetCurvatureStepSize() });
    }
    // set_curve needs a curve; the chord does (only its boundary is read)
    std::vector<Vec3> line(N + 1);
    for (uint32_t i = 0; i <= N; ++i) line[i] = bc.m_startPos + (bc.m_endPos - bc.m_startPos) * ((float)i / (float)N);
    PerturbUtils::BoundaryConditions b = bc;
    b.arclength = maxDs * (float)N;
    twisty_hip_ctx *ctx = nullptr;
    if (twisty_hip_create(ep.device, &ctx) != TWISTY_HIP_OK) {
        error = "twisty_hip_create failed (device " + std::to_string(ep.device) + ")";
        return std::nullopt;
    }
    struct Guard {
        twisty_hip_ctx *c;
        ~Guard() { twisty_hip_destroy(c); }
    } guard { ctx };
    const auto bc13 = PerturbUtils::ToCudaSafe(b);
    int rc = twisty_hip_set_curve(ctx, reinterpret_cast<const float *>(line.data()), N, bc13.data());
    if (rc == TWISTY_HIP_OK)
        rc = twisty_hip_set_multi_tables(ctx, data.data(), bins, (uint32_t)T, meta.data(),
              ep.weightingParameters.absorption, N);
    twisty_hip_partials part {};
    double mn = INFINITY;
    if (rc == TWISTY_HIP_OK)
        rc = twisty_hip_run_generated_multi(ctx, numExperimentPaths, 0, seed, minDs, maxDs, &part, &mn, nullptr, nullptr);
    if (rc != TWISTY_HIP_OK) {
        error = twisty_hip_last_error(ctx);
        return std::nullopt;
    }
    float fwdMs = 0.0f, totMs = 0.0f;
    twisty_hip_last_timing(ctx, &fwdMs, &totMs);
    Result r;
    r.numValidPaths = part.num_valid;
    r.numPathsTotal = numExperimentPaths;
    r.totalWeight.log10 = part.num_valid
          ? part.log10_sum_w + pathNormalizerLog10 - std::log10((double)part.num_valid)
          : -INFINITY;
    r.minPathWeightLog10 = mn + pathNormalizerLog10;
    r.maxPathWeightLog10 = part.max_log10_w + pathNormalizerLog10;
    r.kernelMs = totMs;
    return r;
}

std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &bc, const ExperimentRunner::ExperimentParameters &ep,
      double pathNormalizerLog10, const PathWeighting::CachedMultiArclengthWeightLookupTable &envTables,
      const PathWeighting::CachedMultiArclengthWeightLookupTable & /*objectTables*/, float maxDs, std::string &error,
      uint64_t chainLength)
{
    error.clear();
    const uint32_t N = ep.numSegmentsPerCurve;
    if (N < 4 || numExperimentPaths == 0 || envTables.Count() == 0 || chainLength == 0) {
        error = "need numSegments >= 4, at least one path, one cached table and chainLength >= 1";
        return std::nullopt;
    }
    const float minDs = PathGeneration::CalculateMinimumArclength(bc, N) / (float)N;
    if (maxDs < minDs) return Result {};  // "No solution for current environment"
    const size_t T = envTables.Count();
    const uint32_t bins = (uint32_t)envTables.Table(0).AccessLookupTable().size();
    std::vector<float> data, meta;
    data.reserve(T * bins);
    for (size_t t = 0; t < T; ++t) {
        const auto &tb = envTables.Table(t);
        const auto &v = tb.AccessLookupTable();
        data.insert(data.end(), v.begin(), v.end());
        meta.insert(meta.end(), { tb.GetDs(), tb.GetMinCurvature(), tb.GetMaxCurvature(), tb.GetCurvatureStepSize() });
    }
    std::vector<Vec3> line(N + 1);
    for (uint32_t i = 0; i <= N; ++i) line[i] = bc.m_startPos + (bc.m_endPos - bc.m_startPos) * ((float)i / (float)N);
    PerturbUtils::BoundaryConditions b = bc;
    b.arclength = (minDs + maxDs) * (float)N;
    twisty_hip_ctx *ctx = nullptr;
    if (twisty_hip_create(ep.device, &ctx) != TWISTY_HIP_OK) {
        error = "twisty_hip_create failed (device " + std::to_string(ep.device) + ")";
        return std::nullopt;
    }
    struct Guard {
        twisty_hip_ctx *c;
        ~Guard() { twisty_hip_destroy(c); }
    } guard { ctx };
    // the rate of the ds draw and the loss of every joint (:1663-1664, PathWeighters.cpp:145)
    const float extinction = ep.weightingParameters.absorption + ep.weightingParameters.scatter;
    const uint64_t numChains = (numExperimentPaths + chainLength - 1) / chainLength;
    const auto bc13 = PerturbUtils::ToCudaSafe(b);
    int rc = twisty_hip_set_curve(ctx, reinterpret_cast<const float *>(line.data()), N, bc13.data());
    if (rc == TWISTY_HIP_OK)
        rc = twisty_hip_set_multi_tables(ctx, data.data(), bins, (uint32_t)T, meta.data(), extinction, N);
    twisty_hip_partials part {};
    double mn = INFINITY;
    if (rc == TWISTY_HIP_OK)
        rc = twisty_hip_run_generated_chains(ctx, numChains, chainLength, 0, seed, minDs, maxDs, extinction,
              pathNormalizerLog10, &part, &mn, nullptr, nullptr, nullptr);
    if (rc != TWISTY_HIP_OK) {
        error = twisty_hip_last_error(ctx);
        return std::nullopt;
    }
    float fwdMs = 0.0f, totMs = 0.0f;
    twisty_hip_last_timing(ctx, &fwdMs, &totMs);
    Result r;
    r.numValidPaths = part.num_valid;
    r.numPathsTotal = numChains * chainLength;
    r.totalWeight.log10 = part.num_valid
          ? part.log10_sum_w + pathNormalizerLog10 - std::log10((double)part.num_valid)
          : -INFINITY;
    r.minPathWeightLog10 = mn + pathNormalizerLog10;
    r.maxPathWeightLog10 = part.max_log10_w + pathNormalizerLog10;
    r.kernelMs = totMs;
    return r;
}

}  // namespace ExperimentBase
}  // namespace twisty
