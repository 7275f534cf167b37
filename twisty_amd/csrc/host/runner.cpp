// runner.cpp -- ExperimentRunner and the HIP runner behind it.
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "../../../include/twisty_hip.h"
#include "twisty_host.hpp"

namespace twisty {

std::string BigWeight::ToString(int digits) const
{
    if (std::isinf(log10) && log10 < 0) return "0";
    if (std::isnan(log10)) return "nan";
    const double e = std::floor(log10);
    const double m = std::pow(10.0, log10 - e);
    std::ostringstream os;
    os << std::setprecision(digits) << m << "e" << (e >= 0 ? "+" : "") << (long long)e;
    return os.str();
}

namespace {
const Json &params_of(const Json &cfg) { return cfg.at("experiment").at("experimentParams"); }

std::string time_for_filename()
{
    std::time_t t = std::time(nullptr);
    std::tm tmv {};
    localtime_r(&t, &tmv);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%F_%H-%M-%S", &tmv);
    return buf;
}
}  // namespace

// ExperimentRunner.cpp:48-177 (field names and derived values as there:
// maxBound = 10 / eps, absorption forced to 0 for the simplified model,
// seeds read as ints).
ExperimentRunner::ExperimentParameters ExperimentRunner::ParseExperimentParamsFromConfig(const Json &cfg)
{
    ExperimentParameters ep;
    const Json &p = params_of(cfg);
    ep.numPathsInExperiment = (uint64_t)p.at("pathsToGenerate").number();
    ep.numPathsToSkip = (uint64_t)p.at("pathsToSkip").number();
    ep.experimentName = p.at("name").string();
    ep.experimentDirPath = p.at("experimentDir").string() + "/" + ep.experimentName + "/" + time_for_filename() + "/";
    ep.maxPerturbThreads = (int32_t)p.at("maxPerturbThreads").number();
    ep.maxWeightThreads = (int32_t)p.at("maxWeightThreads").number();
    ep.outputBigFloatWeights = p.at("outputBigFloatWeights").boolean();
    ep.outputPathBatches = p.at("outputPathBatches").boolean();
    ep.useGpu = p.at("useGpu").boolean();
    ep.numSegmentsPerCurve = (uint32_t)p.at("numSegments").number();
    ep.bootstrapSeed = (uint32_t)(int32_t)p.at("random").at("bootstrapSeed").number();
    ep.curvePurturbSeed = (uint32_t)(int32_t)p.at("random").at("perturbSeed").number();
    const Json &w = p.at("weighting");
    ep.weightingParameters.weightingMethod =
          ((int)w.at("weightFunction").number() == 1) ? WeightingMethod::SimplifiedModel : WeightingMethod::RadiativeTransfer;
    const int pm = (int)p.at("perturbMethod").number();
    ep.perturbMethod = pm == 1 ? PerturbMethod::GeometricMinCurvature
                     : pm == 2 ? PerturbMethod::GeometricCombined
                               : PerturbMethod::GeometricRandom;
    ep.weightingParameters.mu = (float)w.at("mu").number();
    ep.weightingParameters.eps = (float)w.at("eps").number();
    ep.weightingParameters.numStepsInt = (uint32_t)w.at("numStepsInt").number();
    ep.weightingParameters.numCurvatureSteps = (uint32_t)w.at("numCurvatureSteps").number();
    ep.weightingParameters.absorption = (float)w.at("absorption").number();
    if (ep.weightingParameters.weightingMethod == WeightingMethod::SimplifiedModel)
        ep.weightingParameters.absorption = 0.0f;
    // The sample configs of the reference list "scatterValues": [s, ...] where
    // its parser reads "scatter"; the first value is taken when "scatter" is absent.
    if (!w.contains("scatter") && w.contains("scatterValues"))
        ep.weightingParameters.scatter = (float)w.at("scatterValues").at(0).number();
    else
        ep.weightingParameters.scatter = (float)w.at("scatter").number();
    if (w.contains("bias")) ep.weightingParameters.bias = (float)w.at("bias").number();
    ep.weightingParameters.minBound = 0.0f;
    ep.weightingParameters.maxBound = 10.0f / ep.weightingParameters.eps;
    if (p.contains("twisty_hip")) {
        const Json &h = p.at("twisty_hip");
        if (h.contains("perturbationsPerPath")) ep.perturbationsPerPath = (uint32_t)h.at("perturbationsPerPath").number();
        if (h.contains("chainLength")) ep.chainLength = (uint32_t)h.at("chainLength").number();
        if (h.contains("device")) ep.device = (int)h.at("device").number();
        if (h.contains("writeFiles")) ep.writeFiles = h.at("writeFiles").boolean();
        if (h.contains("autotune")) ep.autotune = h.at("autotune").boolean();
        if (h.contains("adjointRoulette")) ep.adjointRoulette = (float)h.at("adjointRoulette").number();
    }
    return ep;
}

// Experiments/FullExperiment.cpp:162-191 (directions normalised there).
PerturbUtils::BoundaryConditions ParseBasicExperiment(const Json &cfg)
{
    const Json &b = cfg.at("experiment").at("basicExperiment");
    auto v3 = [&](const char *k) {
        const Json &a = b.at(k);
        return Vec3((float)a.at(0).number(), (float)a.at(1).number(), (float)a.at(2).number());
    };
    PerturbUtils::BoundaryConditions bc;
    bc.m_startPos = v3("startPos");
    bc.m_startDir = v3("startDir").Normalized();
    bc.m_endPos = v3("endPos");
    bc.m_endDir = v3("endDir").Normalized();
    bc.arclength = (float)b.at("arclength").number();
    return bc;
}

ExperimentRunner::ExperimentRunner(ExperimentParameters &experimentParams)
    : m_experimentParams(experimentParams), m_experimentDirPath(experimentParams.experimentDirPath)
{
    if (!experimentParams.perExperimentDirSubfolder.empty())
        m_experimentDirPath += "/" + experimentParams.perExperimentDirSubfolder;
    if (experimentParams.writeFiles) std::filesystem::create_directories(m_experimentDirPath);
}

ExperimentRunner::~ExperimentRunner() = default;

// ExperimentRunner.cpp:196-301 without the path-batch export.
std::optional<ExperimentRunner::ExperimentResults> ExperimentRunner::RunExperiment(
      PerturbUtils::BoundaryConditions boundaryConditions)
{
    const auto t0 = std::chrono::high_resolution_clock::now();
    RunnerSpecificResults r = RunnerSpecificRunExperiment(boundaryConditions);
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (!r.experimentResults) return std::nullopt;
    r.experimentResults->totalExperimentMs = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    r.experimentResults->setupExperimentMs = r.setupMs;
    r.experimentResults->perturbExperimentMs = r.runExperimentMs;
    r.experimentResults->weightingExperimentMs = r.weightingMs;
    return r.experimentResults;
}

FullExperimentRunnerOptimalPerturb_HIP::FullExperimentRunnerOptimalPerturb_HIP(ExperimentParameters &ep)
    : ExperimentRunner(ep)
{
}

FullExperimentRunnerOptimalPerturb_HIP::~FullExperimentRunnerOptimalPerturb_HIP() = default;

namespace {
double log10_add(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double m = std::max(a, b);
    return m + std::log10(std::pow(10.0, a - m) + std::pow(10.0, b - m));
}

// ConvergenceWeights/ (FullExperimentRunnerOptimalPerturb.cpp:90-109, 342-367;
// ExperimentRunner.cpp:399-422): the paths are summed per accumulator of 1e6
// consecutive path ids, the accumulators of every dispatch of 2e10 paths are
// dealt to T "threads" in blocks of ceil(accumulators / T) as the reference's
// dispatch does (:221-266), and after every
// accumulator a line "paths, weight / paths" goes to the thread's
// ConvergenceWeights_<t>.txt and, with outputBigFloatWeights, to the global
// ConvergenceWeights.txt (paths = valid paths so far, weight = sum of
// 10^(w + normaliser) so far).  Fed the per-path log10 weights in path order.
class ConvergenceWriter {
   public:
    ConvergenceWriter(const std::string &experimentDir, uint64_t numPaths, uint32_t threads, double normLog10,
          bool global)
        : m_norm(normLog10)
    {
        const std::filesystem::path dir = std::filesystem::path(experimentDir) / "ConvergenceWeights";
        std::filesystem::create_directories(dir);
        m_threads = threads ? threads : 1;
        m_pathsLeft = numPaths;
        StartDispatch();
        m_files.resize(m_threads);
        for (uint32_t t = 0; t < m_threads; ++t) {
            m_files[t].open((dir / ("ConvergenceWeights_" + std::to_string(t) + ".txt")).string());
            if (!m_files[t].is_open()) throw std::runtime_error("Failed to open ConvergenceWeights_" + std::to_string(t));
        }
        m_threadPaths.assign(m_threads, 0);
        m_threadW.assign(m_threads, -INFINITY);
        if (global) {
            m_global.open((dir / "ConvergenceWeights.txt").string());
            if (!m_global.is_open()) throw std::runtime_error("Failed to open weight convergence file");
        }
    }
    void Add(const float *log10w, uint64_t n)
    {
        for (uint64_t q = 0; q < n; ++q) {
            const float w = log10w[q];
            if (!std::isnan(w)) {  // invalid paths add nothing (:547-549)
                // running max m and sum of 10^(w - m): one pow per path
                if (m_accValid == 0) {
                    m_accMax = w;
                    m_accSum = 1.0;
                } else if (w > m_accMax) {
                    m_accSum = m_accSum * std::pow(10.0, m_accMax - (double)w) + 1.0;
                    m_accMax = w;
                } else {
                    m_accSum += std::pow(10.0, (double)w - m_accMax);
                }
                ++m_accValid;
            }
            ++m_accPaths;
            if (m_accPaths == kAccPaths || m_accPaths == m_dispatchLeft) Close();
        }
    }
    void Finish()
    {
        if (m_accPaths) Close();
    }

   private:
    static constexpr uint64_t kAccPaths = 1000000;       // MaxNumPathsPerCombinedWeight
    static constexpr uint64_t kDispatchPaths = 20000000000ull;  // MaxNumberOfPathsInDispatch
    void StartDispatch()
    {
        m_dispatchLeft = std::min(m_pathsLeft, kDispatchPaths);
        m_pathsLeft -= m_dispatchLeft;
        const uint64_t numAcc = (m_dispatchLeft + kAccPaths - 1) / kAccPaths;
        m_perThread = std::max<uint64_t>(1, (numAcc + m_threads - 1) / m_threads);
        m_acc = 0;
    }
    void Close()
    {
        const double accW = m_accValid ? m_accMax + std::log10(m_accSum) + m_norm : -INFINITY;
        const uint32_t t = (uint32_t)std::min<uint64_t>(m_acc / m_perThread, m_threads - 1);
        m_threadPaths[t] += m_accValid;
        m_threadW[t] = log10_add(m_threadW[t], accW);
        m_files[t] << m_threadPaths[t] << ", " << (BigWeight { m_threadW[t] } / (double)m_threadPaths[t]).ToString()
                   << '\n';
        if (m_global.is_open()) {
            m_globalPaths += m_accValid;
            m_globalW = log10_add(m_globalW, accW);
            m_global << m_globalPaths << ", " << (BigWeight { m_globalW } / (double)m_globalPaths).ToString() << '\n';
        }
        ++m_acc;
        m_dispatchLeft -= m_accPaths;
        m_accPaths = m_accValid = 0;
        if (m_dispatchLeft == 0 && m_pathsLeft) StartDispatch();
    }
    double m_norm;
    uint32_t m_threads = 1;
    uint64_t m_perThread = 1, m_acc = 0, m_accPaths = 0, m_accValid = 0, m_globalPaths = 0;
    uint64_t m_pathsLeft = 0, m_dispatchLeft = 0;
    double m_accMax = 0.0, m_accSum = 0.0, m_globalW = -INFINITY;
    std::vector<std::ofstream> m_files;
    std::ofstream m_global;
    std::vector<uint64_t> m_threadPaths;
    std::vector<double> m_threadW;
};

// Partials of two disjoint path ranges of one experiment.
twisty_hip_partials merge_partials(const twisty_hip_partials &a, const twisty_hip_partials &b)
{
    twisty_hip_partials r;
    r.log10_sum_w = log10_add(a.log10_sum_w, b.log10_sum_w);
    r.log10_sum_w2 = log10_add(a.log10_sum_w2, b.log10_sum_w2);
    r.max_log10_w = std::max(a.max_log10_w, b.max_log10_w);
    r.num_paths = a.num_paths + b.num_paths;
    r.num_valid = a.num_valid + b.num_valid;
    r.num_invalid = a.num_invalid + b.num_invalid;
    r.num_clamped = a.num_clamped + b.num_clamped;
    return r;
}
}  // namespace

// The body of FullExperimentRunnerOptimalPerturb::RunnerSpecificRunExperiment
// (FullExperimentRunnerOptimalPerturb.cpp:48-435) with the per-thread chains
// replaced by one twisty_hip_run: build the weight table (initialised, unlike
// :59-69 which never calls Initialize), the initial curve and the path
// normaliser, then run every path on the device.
ExperimentRunner::RunnerSpecificResults FullExperimentRunnerOptimalPerturb_HIP::RunnerSpecificRunExperiment(
      PerturbUtils::BoundaryConditions bc)
{
    RunnerSpecificResults out;
    m_lastError.clear();
    auto fail = [&](const std::string &msg) {
        m_lastError = msg;
        std::cerr << "FullExperimentRunnerOptimalPerturb_HIP: " << msg << std::endl;
        return out;
    };
    ExperimentParameters &ep = m_experimentParams;
    const auto setupStart = std::chrono::high_resolution_clock::now();
    if (!(bc.arclength > 0.0f)) return fail("arclength must be positive");
    if (ep.weightingParameters.weightingMethod != WeightingMethod::RadiativeTransfer)
        return fail("only the radiative-transfer weighting (weightFunction 0) runs on the HIP kernels");
    if (ep.perturbMethod != PerturbMethod::GeometricRandom)
        return fail("Invalid Perturb Method");  // FullExperimentRunnerOptimalPerturb.cpp:320-325
    if (ep.numPathsToSkip > (1u << 24)) return fail("pathsToSkip above 2^24");
    const uint32_t N = ep.numSegmentsPerCurve;
    if (N < 4) return fail("numSegments must be >= 4");
    const float ds = bc.arclength / (float)N;

    PathWeighting::WeightLookupTableIntegral table(ep.weightingParameters, ds);
    try {
        table.Initialize(ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0);
        if (ep.writeFiles) table.ExportValues(m_experimentDirPath);
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    std::vector<Vec3> positions;
    if (!PathGeneration::InitialArc(bc, N, positions))
        return fail("no initial curve: arclength " + std::to_string(bc.arclength) + " is below the minimum " +
                    std::to_string(PathGeneration::CalculateMinimumArclength(bc, N)));
    const double normLog10 = PathWeighting::NormLog10((int)N, ds, bc);

    twisty_hip_ctx *ctx = nullptr;
    int rc = twisty_hip_create(ep.device, &ctx);
    if (rc != TWISTY_HIP_OK) return fail("twisty_hip_create failed (device " + std::to_string(ep.device) + ")");
    struct Guard {
        twisty_hip_ctx *c;
        ~Guard() { twisty_hip_destroy(c); }
    } guard { ctx };
    const auto &tv = table.AccessLookupTable();
    rc = twisty_hip_set_tables(ctx, tv.data(), (uint32_t)tv.size(), 1, table.GetMinCurvature(),
          table.GetMaxCurvature(), table.GetCurvatureStepSize(), table.GetDs(), ep.weightingParameters.absorption);
    if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    // pathsToSkip: burn-in at the start of every chain, as the reference skips
    // numPathsToSkip paths at the start of every thread's chain
    // (FullExperimentRunnerOptimalPerturb.cpp:286, 469-478, 559-561).
    rc = twisty_hip_set_burn_in(ctx, (uint32_t)ep.numPathsToSkip);
    if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    const auto bc13 = PerturbUtils::ToCudaSafe(bc);
    rc = twisty_hip_set_curve(ctx, reinterpret_cast<const float *>(positions.data()), N, bc13.data());
    if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    const auto setupEnd = std::chrono::high_resolution_clock::now();

    // Seed 0 means "from the clock" (ExperimentRunner.cpp:40-46).
    uint64_t seed = ep.curvePurturbSeed;
    if (seed == 0) seed = (uint64_t)std::time(nullptr);
    const uint64_t nPaths = ep.numPathsInExperiment;
    // Chain length 0: as many chains as 65536 groups can run side by side.
    uint32_t L = ep.chainLength;
    if (L == 0) L = (uint32_t)std::max<uint64_t>(1, (nPaths + 65535) / 65536);
    // The reference always keeps its weight-magnitude histogram
    // (FullExperimentRunnerOptimalPerturb.cpp:26-37: 500 linear bins on
    // [4.69653e-16, 1.17517e-13]) and writes histogram.txt (:394-422).
    constexpr uint32_t kHistBins = 500;
    constexpr double kHistMin = 4.69653e-16, kHistMax = 1.17517e-13;
    if (ep.writeFiles) {
        rc = twisty_hip_set_weight_histogram(ctx, kHistBins, kHistMin, kHistMax, normLog10);
        if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
    }
    twisty_hip_partials part {};
    std::vector<uint64_t> counts(kHistBins, 0);
    float totMs = 0.0f;  // kernel time summed over the runs
    auto addTiming = [&]() {
        float f = 0.0f, t = 0.0f;
        twisty_hip_last_timing(ctx, &f, &t);
        totMs += t;
    };
    // the reference's thread count: maxPerturbThreads, 0 or above the
    // hardware's = hardware_concurrency (FullExperimentRunnerOptimalPerturb.cpp:73-85)
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t convThreads =
          (ep.maxPerturbThreads <= 0 || (uint32_t)ep.maxPerturbThreads > hw) ? hw : (uint32_t)ep.maxPerturbThreads;
    std::unique_ptr<ConvergenceWriter> conv;
    if (ep.writeFiles && nPaths > 0) {
        try {
            conv = std::make_unique<ConvergenceWriter>(m_experimentDirPath, nPaths, convThreads, normLog10,
                  ep.outputBigFloatWeights);
        } catch (const std::exception &e) {
            return fail(e.what());
        }
    }
    if (nPaths > 0 && !ep.outputPathBatches) {
        // One run, or with the convergence files chunks of whole chains (the
        // draws of one run) of about 2^26 paths whose per-path weights come
        // back to the host.
        const uint64_t chunk = conv ? std::max<uint64_t>(1, (uint64_t(1) << 26) / L) * L : nPaths;
        std::vector<float> perPath(conv ? std::min(chunk, nPaths) : 0);
        std::vector<uint64_t> chunkCounts(kHistBins);
        for (uint64_t off = 0; off < nPaths; off += chunk) {
            const uint64_t n = std::min(chunk, nPaths - off);
            twisty_hip_partials c {};
            rc = twisty_hip_run(ctx, n, off, ep.perturbationsPerPath, L, seed, &c, conv ? perPath.data() : nullptr,
                  nullptr);
            if (rc == TWISTY_HIP_OK && ep.writeFiles) rc = twisty_hip_get_weight_histogram(ctx, chunkCounts.data());
            if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
            addTiming();
            if (ep.writeFiles)
                for (uint32_t i = 0; i < kHistBins; ++i) counts[i] += chunkCounts[i];
            if (conv) conv->Add(perPath.data(), n);
            part = off == 0 ? c : merge_partials(part, c);
        }
        if (conv) conv->Finish();
    } else if (nPaths > 0) {
        // Path-batch export: chunks of whole chains (pathOffset a multiple of
        // L, so the draws are those of one run), final points and per-path
        // weights copied back per chunk and appended to the batch files.
        try {
            PathBatchWriter writer(m_experimentDirPath, ep.pathBatchPrepend, ep.experimentName, bc, N);
            rc = twisty_hip_set_keep_final_state(ctx, 1);
            if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
            const uint64_t chunk = std::max<uint64_t>(1, (uint64_t(1) << 18) / L) * L;
            std::vector<float> log10w, points;
            std::vector<uint64_t> chunkCounts(kHistBins);
            bool first = true;
            for (uint64_t off = 0; off < nPaths; off += chunk) {
                const uint64_t n = std::min(chunk, nPaths - off);
                log10w.resize(n);
                points.resize(n * 3ull * (N + 1));
                twisty_hip_partials c {};
                rc = twisty_hip_run(ctx, n, off, ep.perturbationsPerPath, L, seed, &c, log10w.data(), nullptr);
                if (rc == TWISTY_HIP_OK) rc = twisty_hip_get_final_positions(ctx, points.data(), n);
                if (rc == TWISTY_HIP_OK && ep.writeFiles) rc = twisty_hip_get_weight_histogram(ctx, chunkCounts.data());
                if (rc != TWISTY_HIP_OK) return fail(twisty_hip_last_error(ctx));
                addTiming();
                if (ep.writeFiles)
                    for (uint32_t i = 0; i < kHistBins; ++i) counts[i] += chunkCounts[i];
                writer.Add(points.data(), log10w.data(), n, normLog10);
                if (conv) conv->Add(log10w.data(), n);
                part = first ? c : merge_partials(part, c);
                first = false;
            }
            writer.Finish();
            if (conv) conv->Finish();
        } catch (const std::exception &e) {
            return fail(e.what());
        }
    }
    if (nPaths > 0) {
        if (ep.writeFiles) {
            std::ofstream hf(m_experimentDirPath + "/histogram.txt");
            if (!hf.is_open()) return fail("cannot write histogram.txt in " + m_experimentDirPath);
            hf << "Histogram" << '\n';
            for (uint32_t i = 0; i < kHistBins; ++i) {
                const double binMin = kHistMin + (kHistMax - kHistMin) * i / kHistBins;
                const double binMax = kHistMin + (kHistMax - kHistMin) * (i + 1) / kHistBins;
                hf << binMin << " " << binMax << " " << (double)counts[i] << '\n';
            }
        }
    }
    const auto runEnd = std::chrono::high_resolution_clock::now();

    ExperimentResults res;
    res.experimentWeight.log10 = part.log10_sum_w + normLog10;
    res.totalPathsGenerated = nPaths;
    res.numFailedPaths = part.num_invalid;
    res.log10SumSquares = part.log10_sum_w2 + 2.0 * normLog10;
    res.kernelMs = totMs;
    res.numClamped = part.num_clamped;
    out.experimentResults = res;
    out.setupMs = std::chrono::duration_cast<std::chrono::milliseconds>(setupEnd - setupStart).count();
    out.runExperimentMs = std::chrono::duration_cast<std::chrono::milliseconds>(runEnd - setupEnd).count();
    out.weightingMs = 0;
    return out;
}

}  // namespace twisty
