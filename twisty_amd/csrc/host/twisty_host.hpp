// twisty_host.hpp -- C++ host layer: Twisty's runner / weighting / config API
// in front of libtwisty_hip.so.
//
// Class and field names follow the reference so a driver written against
// Twisty/ExperimentRunner.h, PathWeightUtils.h and CurvePerturbUtils.h reads
// the same.  Differences, each forced by a dependency the reference fetches at
// configure time (FMath, Boost.Multiprecision, nlohmann::json):
//   * Farlor::Vector3 -> twisty::Vec3 (three floats, same layout);
//   * boost cpp_dec_float_100 experimentWeight -> twisty::BigWeight, a base-10
//     logarithm in double with the same printed form;
//   * nlohmann::json -> twisty::Json, a small reader of the same documents.
#pragma once

#include <array>
#include <fstream>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

struct twisty_hip_ctx;

namespace twisty {

// ---------------------------------------------------------------------------
// Small vector (Farlor::Vector3 stand-in: three floats).
// ---------------------------------------------------------------------------
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3() = default;
    Vec3(float a, float b, float c) : x(a), y(b), z(c) { }
    Vec3 operator+(const Vec3 &o) const { return { x + o.x, y + o.y, z + o.z }; }
    Vec3 operator-(const Vec3 &o) const { return { x - o.x, y - o.y, z - o.z }; }
    Vec3 operator*(float s) const { return { x * s, y * s, z * s }; }
    float Dot(const Vec3 &o) const { return x * o.x + y * o.y + z * o.z; }
    float Magnitude() const { return std::sqrt(Dot(*this)); }
    Vec3 Normalized() const
    {
        const float m = Magnitude();
        return { x / m, y / m, z / m };
    }
};

namespace PerturbUtils {
// CurvePerturbUtils.h:62-68.
struct BoundaryConditions {
    Vec3 m_startPos = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 m_startDir = Vec3(1.0f, 0.0f, 0.0f);
    Vec3 m_endPos = Vec3(0.0f, 0.0f, 0.0f);
    Vec3 m_endDir = Vec3(1.0f, 0.0f, 0.0f);
    float arclength = 0.0f;
};
// BoundaryConditions_CudaSafe layout (CurvePerturbUtils.h:70-76).
std::array<float, 13> ToCudaSafe(const BoundaryConditions &bc);
}  // namespace PerturbUtils

// ---------------------------------------------------------------------------
// Weighting (PathWeightUtils.h:33-107).
// ---------------------------------------------------------------------------
enum class WeightingMethod : int32_t { RadiativeTransfer = 0, SimplifiedModel = 1 };

struct WeightingParameters {
    float mu = 0.1f;
    uint32_t numStepsInt = 2000;
    float minBound = 0.0f;
    float maxBound = 100.0f;
    float eps = 0.01f;
    float bias = 1.0f;
    float scatter = 0.0f;
    float absorption = 0.0f;
    uint32_t numCurvatureSteps = 10000;
    WeightingMethod weightingMethod = WeightingMethod::RadiativeTransfer;
    // Same field order and hex encoding as the reference's UUID string
    // (PathWeightUtils.h:50-106); the hash is FNV-1a, not std::hash.
    std::pair<std::string, uint64_t> GenerateStringUUID() const;
};

namespace PathWeighting {

float GaussianPhase(float evalLocation, float mu);
double IntegrandRT(double p, double kds, double bds, double eps, double mu);

// PathWeightUtils.h:118-165.
class BaseWeightLookupTable {
   public:
    BaseWeightLookupTable(const WeightingParameters &wp, float ds, float minCurvature, float maxCurvature);
    virtual ~BaseWeightLookupTable() = default;
    virtual void Initialize() = 0;

    const std::vector<float> &AccessLookupTable() const { return m_lookupTable; }
    float GetMinSegmentWeight() const { return m_minSegmentWeight; }
    float GetMaxSegmentWeight() const { return m_maxSegmentWeight; }
    float GetMinCurvature() const { return m_minCurvature; }
    float GetMaxCurvature() const { return m_maxCurvature; }
    float GetDs() const { return m_ds; }
    float GetCurvatureStepSize() const { return m_curvatureStepSize; }
    WeightingParameters GetWeightingParams() const { return m_weightingParams; }

    // Binary cache record: table floats then min/max segment weight
    // (PathWeightUtils.cpp:90-107).
    void StreamOut(std::ostream &os) const;
    void InitializeFromStream(std::istream &is);
    // "curvature, weight" CSV named ds_0x<hex ds>_<hash>_<ExportFilename()>
    // (PathWeightUtils.cpp:55-88).
    void ExportValues(const std::string &directory) const;
    virtual std::string ExportFilename() const = 0;
    std::string UUID() const;

   protected:
    WeightingParameters m_weightingParams;
    float m_ds = 0.0f;
    float m_minCurvature = 0.0f;
    float m_maxCurvature = 0.0f;
    float m_curvatureStepSize = 0.0f;
    float m_minSegmentWeight = 0.0f;
    float m_maxSegmentWeight = 0.0f;
    std::vector<float> m_lookupTable;
};

// WeightLookupTableIntegral (PathWeightUtils.h:195-207, .cpp:281-433): FP64
// Simpson integral of the radiative-transfer integrand per curvature bin.
class WeightLookupTableIntegral : public BaseWeightLookupTable {
   public:
    WeightLookupTableIntegral(const WeightingParameters &wp, float ds);
    void Initialize() override;  // all hardware threads
    void Initialize(int threads);
    std::string ExportFilename() const override { return "RT.csv"; }
    static float Integrate(float curvature, const WeightingParameters &wp, float ds, std::vector<double> &scratch);
    // d T[i] / d eps for every bin, eps the config's `eps` slot (the Gaussian
    // phase width after the reference's argument swap), FP64.  When
    // maxBoundFollowsEps, maxBound = 10/eps moves with eps
    // (ExperimentRunner.cpp:168-171) and so do the Simpson nodes.
    static std::vector<double> EpsDerivative(const WeightingParameters &wp, float ds, bool maxBoundFollowsEps,
          int threads);
};

// CachedMultiArclengthWeightLookupTable (PathWeightUtils.h:168-193,
// .cpp:123-278): numDsSteps radiative-transfer tables for ds on the float grid
// minDs + i (maxDs - minDs)/(numDsSteps - 1), loaded from
// <cacheRoot>/CachedWeightTables/<hash>.cwt when that file exists, otherwise
// computed, written there (the tables' StreamOut records back to back) and
// exported as CSV under CachedWeightTables/<hash>/.  The hash is FNV-1a of the
// reference's UUID string (the reference uses std::hash).  cacheRoot "" is the
// current directory, as the reference.
class CachedMultiArclengthWeightLookupTable {
   public:
    CachedMultiArclengthWeightLookupTable(const WeightingParameters &wp, float minDs, float maxDs, uint32_t numDsSteps,
          const std::string &cacheRoot = "", int threads = 0);
    // The table whose ds is closest to ds (first on ties).
    const BaseWeightLookupTable *GetWeightLookupTable(float ds) const;
    size_t ClosestIndex(float ds) const;
    std::pair<std::string, uint64_t> GetUUID() const { return m_uuid; }
    bool LoadedFromCache() const { return m_loaded; }
    const std::string &CachePath() const { return m_cachePath; }
    size_t Count() const { return m_tables.size(); }
    const BaseWeightLookupTable &Table(size_t i) const { return *m_tables[i]; }

   private:
    float DsAt(uint32_t i) const;
    WeightingParameters m_wp;
    float m_minDs, m_maxDs;
    uint32_t m_numDsSteps;
    std::pair<std::string, uint64_t> m_uuid;
    std::string m_cachePath;
    bool m_loaded = false;
    std::vector<std::unique_ptr<BaseWeightLookupTable>> m_tables;
};

// SimpleWeightLookupTable (PathWeightUtils.h:210-218, .cpp:436-485).
class SimpleWeightLookupTable : public BaseWeightLookupTable {
   public:
    SimpleWeightLookupTable(const WeightingParameters &wp, float ds);
    void Initialize() override;
    std::string ExportFilename() const override { return "Simple.csv"; }
};

// NormalizerStuff::Norm (PathWeightUtils.cpp:721-752): K4 for N = 4, the nested
// quadrature for N = 5, 6 and 1.0 otherwise.  outerSteps / innerSteps are the
// node counts of the K(N) and F(N, M) rules; the defaults are the reference's
// (PathWeightUtils.cpp:694, 665).
double NormLog10(int numberOfSegments, float ds, const PerturbUtils::BoundaryConditions &bc,
      uint32_t outerSteps = 10000, uint32_t innerSteps = 1000);

}  // namespace PathWeighting

namespace PathGeneration {
// PathGeneration.cpp:12-43.
float CalculateMinimumArclength(const PerturbUtils::BoundaryConditions &bc, uint32_t numSegmentsPerCurve);
// Initial curve: hinges of Curve::SetBoundaryConditions (Curve.cpp:20-30) and
// a planar circular arc of N-2 chords of length ds between them.  Returns false
// when the boundary conditions admit no such arc (L below the minimum).
bool InitialArc(const PerturbUtils::BoundaryConditions &bc, uint32_t numSegments, std::vector<Vec3> &positions);
// The same curve in FP64, (N+1)*3 doubles.
bool InitialArcF64(const PerturbUtils::BoundaryConditions &bc, uint32_t numSegments, std::vector<double> &positions);
// d positions / d theta and d phi of the end direction e = (sin theta cos phi,
// sin theta sin phi, cos theta) at the boundary's own end direction, FP64
// (Richardson-extrapolated central differences of InitialArcF64).
bool InitialArcExitJacobian(const PerturbUtils::BoundaryConditions &bc, uint32_t numSegments,
      std::vector<double> &dTheta, std::vector<double> &dPhi);
}  // namespace PathGeneration

// ---------------------------------------------------------------------------
// Path-batch export (outputPathBatches; ExperimentRunner.cpp:210-262,303-396,
// FullExperimentRunnerOptimalPerturb.cpp:599-858), file formats in export.cpp.
// ---------------------------------------------------------------------------
class PathBatchWriter {
   public:
    PathBatchWriter(const std::string &experimentDir, const std::string &prepend, const std::string &experimentName,
          const PerturbUtils::BoundaryConditions &bc, uint32_t numSegments);
    // numPaths paths: (N+1) float3 points each and their log10 weights (NaN =
    // invalid, skipped); the exported weight is 10^(log10 w + normLog10).
    void Add(const float *positions, const float *log10w, uint64_t numPaths, double normLog10);
    void Finish();

   private:
    void Flush();
    uint32_t m_numSegments;
    std::string m_dir, m_experimentName;
    std::ofstream m_pmd, m_pbd, m_bdt, m_pwt, m_fiveFile;
    std::vector<Vec3> m_points, m_fiveCache;
    std::vector<double> m_weights;
    uint64_t m_batchIdx = 0;
};

// ---------------------------------------------------------------------------
// Results (ExperimentRunner.h:73-82).  experimentWeight is carried as log10.
// ---------------------------------------------------------------------------
struct BigWeight {
    double log10 = -INFINITY;
    std::string ToString(int digits = 17) const;  // "m.mmmmeE" scientific form
    BigWeight operator/(double d) const { return { log10 - std::log10(d) }; }
};

class Json;

class ExperimentRunner {
   public:
    enum class PerturbMethod { GeometricRandom, GeometricMinCurvature, GeometricCombined };

    // ExperimentRunner.h:42-68, plus the settings of this implementation that
    // the reference fixes at compile time (read from "twisty_hip" in the
    // config's experimentParams).
    struct ExperimentParameters {
        uint64_t numPathsInExperiment = 0;
        uint64_t numPathsPerBatch = 1000000;
        uint64_t numPathsToSkip = 0;
        uint32_t numSegmentsPerCurve = 4;
        bool allowSaturatePathCount = true;
        float arclength = 0.0f;
        uint32_t exportPathBatchSize = 1000;
        bool exportGeneratedCurves = false;
        std::string experimentName = "Default_Experiment";
        std::string experimentDirPath = "./Default_Experiment/";
        std::string perExperimentDirSubfolder = "";
        std::string pathBatchPrepend = "";
        uint32_t bootstrapSeed = 0;
        uint32_t curvePurturbSeed = 0;
        WeightingParameters weightingParameters;
        PerturbMethod perturbMethod = PerturbMethod::GeometricRandom;
        int32_t maxPerturbThreads = 0;
        int32_t maxWeightThreads = 0;
        bool outputBigFloatWeights = false;
        bool outputPathBatches = false;
        bool useGpu = true;
        // twisty_hip settings
        uint32_t perturbationsPerPath = 1;  // K; the reference applies 1 between weightings
        // Paths sharing one carried state.  The default is the reference CPU
        // runner's chain of 1e6 paths per thread (FullExperimentRunnerOptimalPerturb.cpp:
        // 255-277), so a reference config keeps its Markov-chain statistics;
        // 0 = about 65536 chains (fills the GPU; paths near each chain start
        // stay close to the initial curve unless pathsToSkip burns them in).
        uint32_t chainLength = 1000000;
        int device = 0;
        bool writeFiles = true;             // table CSV, Results-style files
        bool autotune = false;              // batched runs: time the schedules first
        // Gradients: roulette of the reverse sweep (twisty_hip_set_adjoint_roulette;
        // 0 = exact sweep over every path).
        float adjointRoulette = 1.0f;
    };

    struct ExperimentResults {
        BigWeight experimentWeight;
        uint64_t totalPathsGenerated = 0;
        uint64_t numFailedPaths = 0;
        std::chrono::milliseconds::rep totalExperimentMs = 0;
        std::chrono::milliseconds::rep setupExperimentMs = 0;
        std::chrono::milliseconds::rep perturbExperimentMs = 0;
        std::chrono::milliseconds::rep weightingExperimentMs = 0;
        // extras: log10 sum of squared weights, kernel time, clamp count
        double log10SumSquares = -INFINITY;
        double kernelMs = 0.0;
        uint64_t numClamped = 0;
    };

    // ExperimentRunner.cpp:48-177: reads experiment.experimentParams.*.
    // Unlike the reference, a missing or mistyped field is an error (the
    // reference prints and returns defaults).
    static ExperimentParameters ParseExperimentParamsFromConfig(const Json &experimentConfig);

    explicit ExperimentRunner(ExperimentParameters &experimentParams);
    virtual ~ExperimentRunner();
    std::optional<ExperimentResults> RunExperiment(PerturbUtils::BoundaryConditions boundaryConditions);

   protected:
    struct RunnerSpecificResults {
        std::optional<ExperimentResults> experimentResults;
        std::chrono::milliseconds::rep experimentRuntimeTotalMs = 0;
        std::chrono::milliseconds::rep setupMs = 0;
        std::chrono::milliseconds::rep runExperimentMs = 0;
        std::chrono::milliseconds::rep weightingMs = 0;
    };
    virtual RunnerSpecificResults RunnerSpecificRunExperiment(PerturbUtils::BoundaryConditions boundaryConditions) = 0;

    ExperimentParameters &m_experimentParams;
    std::string m_experimentDirPath;
};

// The HIP runner: the drop-in for FullExperimentRunnerOptimalPerturb
// (FullExperimentRunnerOptimalPerturb.h:26-74) and its CUDA variant
// (FullExperimentRunnerOptimalPerturbOptimized_GPU.h).
class FullExperimentRunnerOptimalPerturb_HIP : public ExperimentRunner {
   public:
    explicit FullExperimentRunnerOptimalPerturb_HIP(ExperimentParameters &experimentParams);
    ~FullExperimentRunnerOptimalPerturb_HIP() override;
    // Last error reported by the HIP library (empty when none).
    const std::string &LastError() const { return m_lastError; }

   protected:
    RunnerSpecificResults RunnerSpecificRunExperiment(PerturbUtils::BoundaryConditions boundaryConditions) override;

   private:
    std::string m_lastError;
};

// Many experiments of one parameter set in one persistent batched launch: the
// callers are the multi-experiment drivers (BeamSpreadExperiment.cpp:209-284,
// RingBenchmark.cpp:166-215), which in the reference construct and run one
// ExperimentRunner per geometry.  Every experiment uses the parameters'
// path count, segment count, weighting, K, chain length and burn-in; tables
// are built once per distinct ds.  Results are those RunExperiment gives for
// the same experiment (seed 0 = from the clock).
class BatchedExperimentRunner {
   public:
    struct Experiment {
        PerturbUtils::BoundaryConditions boundaryConditions;
        uint64_t seed = 0;
    };
    explicit BatchedExperimentRunner(const ExperimentRunner::ExperimentParameters &experimentParams);
    // nullopt on failure (LastError() says why).
    std::optional<std::vector<ExperimentRunner::ExperimentResults>> RunExperiments(
          const std::vector<Experiment> &experiments);
    const std::string &LastError() const { return m_lastError; }
    uint32_t NumTables() const { return m_numTables; }
    double DeviceMs() const { return m_deviceMs; }

   private:
    ExperimentRunner::ExperimentParameters m_ep;
    std::string m_lastError;
    uint32_t m_numTables = 0;
    double m_deviceMs = 0.0;
};

// Gradients of one experiment's log-estimator E = log10(sum_p w_p / P) (no
// normaliser; it is 1 for N >= 7, PathWeightUtils.cpp:721-752) through the
// adjoint: dE/dT, dE/d eps (the config's eps slot, maxBound = 10/eps moving
// with it, ExperimentRunner.cpp:168-171) and dE/d(theta, phi) of the end
// direction e = (sin theta cos phi, sin theta sin phi, cos theta), the last two
// chained on the host in FP64 through dT/d eps and the initial arc's Jacobian.
// Paths restart from the initial curve (chain length 1); K from
// perturbationsPerPath; seed 0 = from the clock.  The reference has no
// gradients.
struct ExperimentGradients {
    double log10Estimate = 0.0;
    double dEps = 0.0, dTheta = 0.0, dPhi = 0.0;
    std::vector<double> dTable;
    double kernelMs = 0.0;
    // the exit angles' standard deviation added by the roulette of the
    // reverse sweep (0 for the exact sweep) and their Monte Carlo standard error
    double dThetaRouletteStd = 0.0, dThetaMcStd = 0.0, dPhiRouletteStd = 0.0, dPhiMcStd = 0.0;
    uint64_t pathsSwept = 0;
};
std::optional<ExperimentGradients> ComputeExperimentGradients(const ExperimentRunner::ExperimentParameters &ep,
      const PerturbUtils::BoundaryConditions &bc, std::string &error);

// The independent-path estimator ExperimentBase::MSegmentPathGenerationMC
// (Experiments/ExperimentBase.cpp:1246-1386) on twisty_hip_run_generated:
// every path built from scratch by recursive sphere intersection, weighted
// with the radiative-transfer table of ds = arclength / N.  totalWeight is the
// reference's sum over valid paths of 10^(w + normaliser) / numValidPaths.
namespace ExperimentBase {
struct Result {  // ExperimentBase.h Result
    uint64_t numValidPaths = 0;
    uint64_t numPathsTotal = 0;
    BigWeight totalWeight;
    double minPathWeightLog10 = INFINITY;
    double maxPathWeightLog10 = -INFINITY;
    double kernelMs = 0.0;
};
std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &experimentGeometry,
      const ExperimentRunner::ExperimentParameters &experimentParams, double pathNormalizerLog10,
      std::string &error);
// The multi-arclength overload (ExperimentBase.cpp:1388-1597): ds drawn per path
// on [minDs, maxDs] with minDs = the quadratic minimum arclength of :1437-1462
// over N (an empty Result when maxDs < minDs, :1466-1469), the nearest table of
// the cached set.
std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &experimentGeometry,
      const ExperimentRunner::ExperimentParameters &experimentParams, double pathNormalizerLog10,
      const PathWeighting::CachedMultiArclengthWeightLookupTable &cachedWeightLookupTable, float maxDs,
      std::string &error);
// The two-table overload (ExperimentBase.cpp:1599-1818), the estimator of the
// NoisyCircle experiments, on twisty_hip_run_generated_chains: Metropolis-
// Hastings chains of chainLength proposals (the reference runs one per OpenMP
// thread; here ceil(numExperimentPaths / chainLength) of them, so numPathsTotal
// is that many chains times chainLength), minDs = CalculateMinimumArclength / N
// (an empty Result when maxDs < minDs, :1639-1642), ds on [minDs, minDs +
// maxDs], every joint weighted with the environment table nearest its ds and
// absorption + scatter.  The object tables are taken and not read: the
// reference's weighter has its object branch commented out
// (PathWeighters.cpp:137-141).  totalWeight is the sum over valid proposals of
// 10^(chain weight + normaliser) / numValidPaths.
std::optional<Result> MSegmentPathGenerationMC(uint64_t seed, uint64_t numExperimentPaths,
      const PerturbUtils::BoundaryConditions &experimentGeometry,
      const ExperimentRunner::ExperimentParameters &experimentParams, double pathNormalizerLog10,
      const PathWeighting::CachedMultiArclengthWeightLookupTable &environmentCachedWeightLookupTable,
      const PathWeighting::CachedMultiArclengthWeightLookupTable &objectCachedWeightLookupTable, float maxDs,
      std::string &error, uint64_t chainLength = 64);
// ExperimentBase.cpp:1437-1462: the smaller positive root of the quadratic in
// the arclength for M segments (1e6 when neither root is positive).
float QuadraticMinimumArclength(const PerturbUtils::BoundaryConditions &experimentGeometry, uint32_t M);
}  // namespace ExperimentBase

// ---------------------------------------------------------------------------
// Minimal JSON value (objects, arrays, numbers, strings, booleans, null).
// ---------------------------------------------------------------------------
class Json {
   public:
    enum class Type { Null, Bool, Number, String, Array, Object };
    Json() = default;
    static Json Parse(const std::string &text);  // throws std::runtime_error
    static Json ParseFile(const std::string &path);
    Type type() const { return m_type; }
    bool contains(const std::string &key) const;
    const Json &at(const std::string &key) const;  // throws when missing
    const Json &at(size_t i) const;
    size_t size() const;
    double number() const;
    bool boolean() const;
    const std::string &string() const;

   private:
    Type m_type = Type::Null;
    bool m_bool = false;
    double m_number = 0.0;
    std::string m_string;
    std::vector<Json> m_array;
    std::map<std::string, Json> m_object;
    friend class JsonParser;
};

// BoundaryConditions from experiment.basicExperiment (Experiments/FullExperiment.cpp:162-191).
PerturbUtils::BoundaryConditions ParseBasicExperiment(const Json &experimentConfig);

}  // namespace twisty
