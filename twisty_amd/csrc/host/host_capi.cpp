// host_capi.cpp -- C ABI of libtwisty_host.so (include/twisty_host.h).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../../include/twisty_host.h"
#include "twisty_host.hpp"

using namespace twisty;

namespace {
PerturbUtils::BoundaryConditions from13(const float b[13])
{
    PerturbUtils::BoundaryConditions bc;
    bc.m_startPos = Vec3(b[0], b[1], b[2]);
    bc.m_startDir = Vec3(b[3], b[4], b[5]);
    bc.m_endPos = Vec3(b[6], b[7], b[8]);
    bc.m_endDir = Vec3(b[9], b[10], b[11]);
    bc.arclength = b[12];
    return bc;
}

void put_error(char *buf, uint32_t len, const std::string &msg)
{
    if (!buf || len == 0) return;
    const size_t n = std::min<size_t>(len - 1, msg.size());
    std::memcpy(buf, msg.data(), n);
    buf[n] = '\0';
}
}  // namespace

extern "C" {

int twisty_host_build_rt_table(float mu, float eps, float scatter, float absorption, float bias,
      uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound, float maxBound, float ds, int threads,
      float *tableOut, float *metaOut)
{
    if (!tableOut || !metaOut || numCurvatureSteps < 2 || numStepsInt < 4 || numStepsInt % 2 != 0 || !(ds > 0.0f))
        return 1;
    WeightingParameters wp;
    wp.mu = mu;
    wp.eps = eps;
    wp.scatter = scatter;
    wp.absorption = absorption;
    wp.bias = bias;
    wp.numStepsInt = numStepsInt;
    wp.numCurvatureSteps = numCurvatureSteps;
    wp.minBound = minBound;
    wp.maxBound = maxBound;
    try {
        PathWeighting::WeightLookupTableIntegral t(wp, ds);
        t.Initialize(threads);
        const auto &v = t.AccessLookupTable();
        std::memcpy(tableOut, v.data(), v.size() * sizeof(float));
        const float meta[6] = { t.GetDs(), t.GetMinCurvature(), t.GetMaxCurvature(), t.GetCurvatureStepSize(),
            t.GetMinSegmentWeight(), t.GetMaxSegmentWeight() };
        std::memcpy(metaOut, meta, sizeof(meta));
    } catch (const std::exception &) {
        return 1;
    }
    return 0;
}

int twisty_host_build_rt_tables(uint32_t count, const float *eps, const float *ds, const float *maxBound, float mu,
      float scatter, float absorption, float bias, uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound,
      int threads, float *tablesOut, float *metaOut)
{
    if (!eps || !ds || !tablesOut || !metaOut || numCurvatureSteps < 2 || numStepsInt < 4 || numStepsInt % 2 != 0)
        return 1;
    for (uint32_t k = 0; k < count; ++k)
        if (!(ds[k] > 0.0f) || !(eps[k] > 0.0f)) return 1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<uint32_t>((uint32_t)threads, std::max(1u, count));
    std::atomic<uint32_t> next { 0 };
    std::atomic<int> failed { 0 };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (uint32_t k = next++; k < count; k = next++) {
                WeightingParameters wp;
                wp.mu = mu;
                wp.eps = eps[k];
                wp.scatter = scatter;
                wp.absorption = absorption;
                wp.bias = bias;
                wp.numStepsInt = numStepsInt;
                wp.numCurvatureSteps = numCurvatureSteps;
                wp.minBound = minBound;
                // ExperimentRunner.cpp:168-171: maxBound = 10 / eps
                wp.maxBound = maxBound ? maxBound[k] : 10.0f / eps[k];
                try {
                    PathWeighting::WeightLookupTableIntegral tb(wp, ds[k]);
                    tb.Initialize(1);
                    const auto &v = tb.AccessLookupTable();
                    std::memcpy(tablesOut + (size_t)k * numCurvatureSteps, v.data(), v.size() * sizeof(float));
                    const float meta[6] = { tb.GetDs(), tb.GetMinCurvature(), tb.GetMaxCurvature(),
                        tb.GetCurvatureStepSize(), tb.GetMinSegmentWeight(), tb.GetMaxSegmentWeight() };
                    std::memcpy(metaOut + (size_t)k * 6, meta, sizeof(meta));
                } catch (const std::exception &) {
                    failed = 1;
                }
            }
        });
    for (auto &th : pool) th.join();
    return failed ? 1 : 0;
}

int twisty_host_cached_rt_tables(const char *cacheRoot, float mu, float eps, float scatter, float absorption,
      float bias, uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound, float maxBound, float minDs,
      float maxDs, uint32_t numDsSteps, int threads, float *tablesOut, float *metaOut, int *loadedFromCache,
      char *errorOut, uint32_t errorLen)
{
    if (!tablesOut || !metaOut || numCurvatureSteps < 2 || numStepsInt < 4 || numStepsInt % 2 != 0) return 1;
    WeightingParameters wp;
    wp.mu = mu;
    wp.eps = eps;
    wp.scatter = scatter;
    wp.absorption = absorption;
    wp.bias = bias;
    wp.numStepsInt = numStepsInt;
    wp.numCurvatureSteps = numCurvatureSteps;
    wp.minBound = minBound;
    wp.maxBound = maxBound;
    try {
        PathWeighting::CachedMultiArclengthWeightLookupTable c(wp, minDs, maxDs, numDsSteps,
              cacheRoot ? std::string(cacheRoot) : std::string(), threads);
        for (size_t i = 0; i < c.Count(); ++i) {
            const auto &t = c.Table(i);
            const auto &v = t.AccessLookupTable();
            std::memcpy(tablesOut + i * numCurvatureSteps, v.data(), v.size() * sizeof(float));
            const float meta[6] = { t.GetDs(), t.GetMinCurvature(), t.GetMaxCurvature(), t.GetCurvatureStepSize(),
                t.GetMinSegmentWeight(), t.GetMaxSegmentWeight() };
            std::memcpy(metaOut + 6 * i, meta, sizeof(meta));
        }
        if (loadedFromCache) *loadedFromCache = c.LoadedFromCache() ? 1 : 0;
        put_error(errorOut, errorLen, c.CachePath());
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_rt_table_eps_derivative(float mu, float eps, float scatter, float bias, uint32_t numStepsInt,
      uint32_t numCurvatureSteps, float minBound, float maxBound, int maxBoundFollowsEps, float ds, int threads,
      double *dTableDEps)
{
    if (!dTableDEps || numCurvatureSteps < 2 || numStepsInt < 4 || numStepsInt % 2 != 0 || !(ds > 0.0f) ||
          !(eps > 0.0f))
        return 1;
    WeightingParameters wp;
    wp.mu = mu;
    wp.eps = eps;
    wp.scatter = scatter;
    wp.bias = bias;
    wp.numStepsInt = numStepsInt;
    wp.numCurvatureSteps = numCurvatureSteps;
    wp.minBound = minBound;
    wp.maxBound = maxBound;
    try {
        const std::vector<double> d =
              PathWeighting::WeightLookupTableIntegral::EpsDerivative(wp, ds, maxBoundFollowsEps != 0, threads);
        std::memcpy(dTableDEps, d.data(), d.size() * sizeof(double));
    } catch (const std::exception &) {
        return 1;
    }
    return 0;
}

int twisty_host_initial_arc_exit_jacobian(const float boundary[13], uint32_t numSegments, double *dPosDTheta,
      double *dPosDPhi)
{
    if (!boundary || !dPosDTheta || !dPosDPhi) return 1;
    std::vector<double> a, b;
    if (!PathGeneration::InitialArcExitJacobian(from13(boundary), numSegments, a, b)) return 1;
    std::memcpy(dPosDTheta, a.data(), a.size() * sizeof(double));
    std::memcpy(dPosDPhi, b.data(), b.size() * sizeof(double));
    return 0;
}

int twisty_host_initial_arc(const float boundary[13], uint32_t numSegments, float *positionsOut)
{
    if (!boundary || !positionsOut) return 1;
    std::vector<Vec3> pos;
    if (!PathGeneration::InitialArc(from13(boundary), numSegments, pos)) return 1;
    std::memcpy(positionsOut, pos.data(), pos.size() * sizeof(Vec3));
    return 0;
}

float twisty_host_minimum_arclength(const float boundary[13], uint32_t numSegments)
{
    return PathGeneration::CalculateMinimumArclength(from13(boundary), numSegments);
}

double twisty_host_norm_log10(uint32_t numSegments, float ds, const float boundary[13])
{
    return PathWeighting::NormLog10((int)numSegments, ds, from13(boundary));
}

double twisty_host_norm_log10_steps(uint32_t numSegments, float ds, const float boundary[13], uint32_t outerSteps,
      uint32_t innerSteps)
{
    if (outerSteps == 0 || innerSteps == 0) return std::numeric_limits<double>::quiet_NaN();
    return PathWeighting::NormLog10((int)numSegments, ds, from13(boundary), outerSteps, innerSteps);
}

int twisty_host_write_path_batches(const char *experimentDir, const char *prepend, const char *experimentName,
      const float boundary[13], uint32_t numSegments, const float *positions, const float *log10w, uint64_t numPaths,
      double normLog10, char *errorOut, uint32_t errorLen)
{
    if (!experimentDir || !boundary || (numPaths > 0 && (!positions || !log10w))) return 1;
    try {
        PathBatchWriter w(experimentDir, prepend ? prepend : "", experimentName ? experimentName : "Default_Experiment",
              from13(boundary), numSegments);
        w.Add(positions, log10w, numPaths, normLog10);
        w.Finish();
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_run_experiments(const char *jsonText, const float *boundaries, const uint64_t *seeds,
      uint32_t numExperiments, twisty_host_result *out, double *deviceMs, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !boundaries || !seeds || !out || numExperiments == 0) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        const ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        std::vector<BatchedExperimentRunner::Experiment> exps(numExperiments);
        for (uint32_t e = 0; e < numExperiments; ++e) {
            exps[e].boundaryConditions = from13(boundaries + 13ull * e);
            exps[e].seed = seeds[e];
        }
        BatchedExperimentRunner runner(ep);
        const auto res = runner.RunExperiments(exps);
        if (!res) {
            put_error(errorOut, errorLen, runner.LastError());
            return 1;
        }
        for (uint32_t e = 0; e < numExperiments; ++e) {
            const ExperimentRunner::ExperimentResults &r = (*res)[e];
            twisty_host_result &o = out[e];
            o.log10_experiment_weight = r.experimentWeight.log10;
            o.log10_avg_path_weight = (r.experimentWeight / (double)r.totalPathsGenerated).log10;
            o.log10_sum_squares = r.log10SumSquares;
            o.total_paths = r.totalPathsGenerated;
            o.failed_paths = r.numFailedPaths;
            o.clamped_joints = r.numClamped;
            o.setup_ms = (double)r.setupExperimentMs;
            o.perturb_ms = (double)r.perturbExperimentMs;
            o.weighting_ms = (double)r.weightingExperimentMs;
            o.total_ms = (double)r.totalExperimentMs;
            o.kernel_ms = r.kernelMs;
        }
        if (deviceMs) *deviceMs = runner.DeviceMs();
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_path_generation_mc(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      twisty_host_generation_result *out, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !out) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        const ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        const PerturbUtils::BoundaryConditions bc = ParseBasicExperiment(cfg);
        std::string err;
        const auto r = ExperimentBase::MSegmentPathGenerationMC(seed, ep.numPathsInExperiment, bc, ep,
              pathNormalizerLog10, err);
        if (!r) {
            put_error(errorOut, errorLen, err);
            return 1;
        }
        out->log10_total_weight = r->totalWeight.log10;
        out->min_log10_w = r->minPathWeightLog10;
        out->max_log10_w = r->maxPathWeightLog10;
        out->num_valid = r->numValidPaths;
        out->num_total = r->numPathsTotal;
        out->kernel_ms = r->kernelMs;
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_path_generation_mc_multi(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      float minDsTables, float maxDsTables, uint32_t numDsSteps, float maxDs, const char *cacheRoot,
      twisty_host_generation_result *out, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !out) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        const ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        const PerturbUtils::BoundaryConditions bc = ParseBasicExperiment(cfg);
        const PathWeighting::CachedMultiArclengthWeightLookupTable tables(ep.weightingParameters, minDsTables,
              maxDsTables, numDsSteps, cacheRoot ? cacheRoot : "", ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0);
        std::string err;
        const auto r = ExperimentBase::MSegmentPathGenerationMC(seed, ep.numPathsInExperiment, bc, ep,
              pathNormalizerLog10, tables, maxDs, err);
        if (!r) {
            put_error(errorOut, errorLen, err);
            return 1;
        }
        out->log10_total_weight = r->totalWeight.log10;
        out->min_log10_w = r->minPathWeightLog10;
        out->max_log10_w = r->maxPathWeightLog10;
        out->num_valid = r->numValidPaths;
        out->num_total = r->numPathsTotal;
        out->kernel_ms = r->kernelMs;
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_path_generation_mc_chains(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      float minDsTables, float maxDsTables, uint32_t numDsSteps, float maxDs, uint64_t chainLength,
      const char *cacheRoot, twisty_host_generation_result *out, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !out) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        const ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        const PerturbUtils::BoundaryConditions bc = ParseBasicExperiment(cfg);
        const PathWeighting::CachedMultiArclengthWeightLookupTable tables(ep.weightingParameters, minDsTables,
              maxDsTables, numDsSteps, cacheRoot ? cacheRoot : "", ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0);
        std::string err;
        // the object set is not read (the reference's weighter has its object branch commented out)
        const auto r = ExperimentBase::MSegmentPathGenerationMC(seed, ep.numPathsInExperiment, bc, ep,
              pathNormalizerLog10, tables, tables, maxDs, err, chainLength);
        if (!r) {
            put_error(errorOut, errorLen, err);
            return 1;
        }
        out->log10_total_weight = r->totalWeight.log10;
        out->min_log10_w = r->minPathWeightLog10;
        out->max_log10_w = r->maxPathWeightLog10;
        out->num_valid = r->numValidPaths;
        out->num_total = r->numPathsTotal;
        out->kernel_ms = r->kernelMs;
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_experiment_gradients(const char *jsonText, twisty_host_gradients *out, double *dTableOut,
      uint32_t dTableLen, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !out) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        const ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        const PerturbUtils::BoundaryConditions bc = ParseBasicExperiment(cfg);
        if (dTableOut && dTableLen != ep.weightingParameters.numCurvatureSteps) {
            put_error(errorOut, errorLen, "dTableLen must equal numCurvatureSteps");
            return 1;
        }
        std::string err;
        const auto g = ComputeExperimentGradients(ep, bc, err);
        if (!g) {
            put_error(errorOut, errorLen, err);
            return 1;
        }
        out->log10_estimate = g->log10Estimate;
        out->d_eps = g->dEps;
        out->d_theta = g->dTheta;
        out->d_phi = g->dPhi;
        out->kernel_ms = g->kernelMs;
        out->d_theta_roulette_std = g->dThetaRouletteStd;
        out->d_theta_mc_std = g->dThetaMcStd;
        out->d_phi_roulette_std = g->dPhiRouletteStd;
        out->d_phi_mc_std = g->dPhiMcStd;
        out->paths_swept = g->pathsSwept;
        if (dTableOut) std::copy(g->dTable.begin(), g->dTable.end(), dTableOut);
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

int twisty_host_run_config(const char *jsonText, twisty_host_result *out, char *errorOut, uint32_t errorLen)
{
    if (!jsonText || !out) return 1;
    try {
        const Json cfg = Json::Parse(jsonText);
        ExperimentRunner::ExperimentParameters ep = ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        PerturbUtils::BoundaryConditions bc = ParseBasicExperiment(cfg);
        ep.arclength = bc.arclength;
        ep.perExperimentDirSubfolder = "main";
        FullExperimentRunnerOptimalPerturb_HIP runner(ep);
        const auto res = runner.RunExperiment(bc);
        if (!res) {
            put_error(errorOut, errorLen, runner.LastError());
            return 1;
        }
        out->log10_experiment_weight = res->experimentWeight.log10;
        out->log10_avg_path_weight = (res->experimentWeight / (double)res->totalPathsGenerated).log10;
        out->log10_sum_squares = res->log10SumSquares;
        out->total_paths = res->totalPathsGenerated;
        out->failed_paths = res->numFailedPaths;
        out->clamped_joints = res->numClamped;
        out->setup_ms = (double)res->setupExperimentMs;
        out->perturb_ms = (double)res->perturbExperimentMs;
        out->weighting_ms = (double)res->weightingExperimentMs;
        out->total_ms = (double)res->totalExperimentMs;
        out->kernel_ms = res->kernelMs;
    } catch (const std::exception &e) {
        put_error(errorOut, errorLen, e.what());
        return 1;
    }
    return 0;
}

}  // extern "C"
