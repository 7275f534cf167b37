/*
 * twisty_host.h -- C ABI of libtwisty_host.so, the C++ host layer (weight
 * tables, initial curve, ExperimentRunner) in front of libtwisty_hip.so.
 *
 * These entry points let non-C++ callers (the Python tooling, ctypes) reach
 * the host layer.  Same conventions as twisty_hip.h: status codes, host
 * pointers, synchronous.
 */
#ifndef TWISTY_HOST_H
#define TWISTY_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ExperimentRunner::ExperimentResults (ExperimentRunner.h:73-82) with the
 * experiment weight as log10 (the reference holds a Boost cpp_dec_float_100). */
typedef struct twisty_host_result {
    double log10_experiment_weight; /* log10 of sum_valid 10^{w_p} * path normaliser */
    double log10_avg_path_weight;   /* log10 of experimentWeight / totalPathsGenerated
                                       (Experiments/FullExperiment.cpp:244) */
    double log10_sum_squares;       /* log10 of sum_valid 10^{2 w_p} * normaliser^2 */
    uint64_t total_paths;
    uint64_t failed_paths;
    uint64_t clamped_joints;
    double setup_ms, perturb_ms, weighting_ms, total_ms, kernel_ms;
} twisty_host_result;

/* WeightLookupTableIntegral(params, ds).Initialize() (PathWeightUtils.cpp:
 * 281-348): numCurvatureSteps floats into tableOut; metaOut receives
 * [ds, minCurvature, maxCurvature, curvatureStepSize, minSegmentWeight,
 * maxSegmentWeight].  threads <= 0 uses every hardware thread. */
int twisty_host_build_rt_table(float mu, float eps, float scatter, float absorption, float bias,
      uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound, float maxBound, float ds, int threads,
      float *tableOut, float *metaOut);

/* Initial curve for BoundaryConditions_CudaSafe `boundary` (13 floats): hinges
 * as Curve::SetBoundaryConditions (Curve.cpp:20-30), circular arc between
 * them; (numSegments+1)*3 floats.  Returns 1 when no arc exists. */
int twisty_host_initial_arc(const float boundary[13], uint32_t numSegments, float *positionsOut);

/* count tables as twisty_host_build_rt_table builds them, table k for its own
 * (eps[k], ds[k]) and maxBound[k] (maxBound NULL: 10 / eps[k] in float, as
 * ExperimentRunner.cpp:168-171 sets it), the other parameters shared.  What
 * FullExperimentRunnerOptimalPerturb.cpp:55-65 does once per experiment, for a
 * batch of experiments: each table is built on one thread, threads tables at a
 * time (<= 0: every hardware thread).  tablesOut: count * numCurvatureSteps
 * floats; metaOut: 6 floats per table. */
int twisty_host_build_rt_tables(uint32_t count, const float *eps, const float *ds, const float *maxBound, float mu,
      float scatter, float absorption, float bias, uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound,
      int threads, float *tablesOut, float *metaOut);

/* CachedMultiArclengthWeightLookupTable (PathWeightUtils.cpp:123-278):
 * numDsSteps tables for ds = minDs + i (maxDs - minDs)/(numDsSteps - 1), read
 * from <cacheRoot>/CachedWeightTables/<hash>.cwt when present, otherwise
 * computed and written there (cacheRoot NULL or "" = current directory).
 * tablesOut: numDsSteps*numCurvatureSteps floats; metaOut: 6 floats per table
 * as twisty_host_build_rt_table; *loadedFromCache = 1 when read from the file.
 * errorOut receives the cache file path on success, the message on failure. */
int twisty_host_cached_rt_tables(const char *cacheRoot, float mu, float eps, float scatter, float absorption,
      float bias, uint32_t numStepsInt, uint32_t numCurvatureSteps, float minBound, float maxBound, float minDs,
      float maxDs, uint32_t numDsSteps, int threads, float *tablesOut, float *metaOut, int *loadedFromCache,
      char *errorOut, uint32_t errorLen);

/* d T[i] / d eps of the table twisty_host_build_rt_table builds, in FP64, for
 * the config's eps slot (the Gaussian phase width the reference actually uses,
 * PathWeightUtils.cpp:401-420 with its (mu, eps) swap).  maxBoundFollowsEps != 0
 * differentiates maxBound = 10/eps as well (ExperimentRunner.cpp:168-171).
 * The reference has no derivative; this feeds the phase-width gradient. */
int twisty_host_rt_table_eps_derivative(float mu, float eps, float scatter, float bias, uint32_t numStepsInt,
      uint32_t numCurvatureSteps, float minBound, float maxBound, int maxBoundFollowsEps, float ds, int threads,
      double *dTableDEps);

/* d positions / d theta and d phi ((numSegments+1)*3 doubles each) of the
 * initial curve of twisty_host_initial_arc when the end direction is
 * (sin theta cos phi, sin theta sin phi, cos theta), at the boundary's own
 * (normalised) end direction.  Feeds the exit-direction gradient. */
int twisty_host_initial_arc_exit_jacobian(const float boundary[13], uint32_t numSegments, double *dPosDTheta,
      double *dPosDPhi);

/* PathGeneration::CalculateMinimumArclength (PathGeneration.cpp:12-43). */
float twisty_host_minimum_arclength(const float boundary[13], uint32_t numSegments);

/* log10 of NormalizerStuff::Norm (PathWeightUtils.cpp:721-752). */
double twisty_host_norm_log10(uint32_t numSegments, float ds, const float boundary[13]);

/* The same with the node counts of the nested N = 5, 6 quadrature as arguments:
 * outerSteps for the K(N) rule, innerSteps for each F(N, M) rule (the reference
 * hard-codes 10000 and 1000, PathWeightUtils.cpp:694, 665; twisty_host_norm_log10
 * uses those).  Either count 0 returns NaN. */
double twisty_host_norm_log10_steps(uint32_t numSegments, float ds, const float boundary[13], uint32_t outerSteps,
      uint32_t innerSteps);

/* Path-batch export of numPaths paths ((numSegments+1)*3 floats each, and their
 * log10 weights, NaN = invalid and skipped) under <experimentDir>/GeneratedCurves:
 * BoundaryConditions.bcf, <prepend>Paths_Metadata.pmd, <prepend>Paths_Binary.pbd,
 * <prepend>Log10PathWeights_Binary.bdt (the weight 10^(log10 w + normLog10), a
 * double per path), <prepend>Log10PathWeights_Text.pwt,
 * <prepend>FiveSegment_Binary.bdt and index.json, in batches of up to 1e6 paths.
 * The writer of ExperimentRunner.cpp:210-262,303-396 and
 * FullExperimentRunnerOptimalPerturb.cpp:599-858; the runner uses it when
 * outputPathBatches is set. */
int twisty_host_write_path_batches(const char *experimentDir, const char *prepend, const char *experimentName,
      const float boundary[13], uint32_t numSegments, const float *positions, const float *log10w, uint64_t numPaths,
      double normLog10, char *errorOut, uint32_t errorLen);

/* BatchedExperimentRunner: numExperiments experiments sharing the
 * experimentParams of a JSON config (text), each with its own boundary
 * (13 floats, BoundaryConditions_CudaSafe layout with the arclength) and seed
 * (0 = from the clock), run in one batched launch; out[e] as
 * twisty_host_run_config would report experiment e run alone with the same
 * chain length.  Replaces the one-runner-per-geometry loops of the
 * multi-experiment drivers (Experiments/BeamSpreadExperiment.cpp:209-284,
 * Experiments/RingBenchmark.cpp:166-215).  deviceMs (nullable) receives the
 * batched launch's device time. */
int twisty_host_run_experiments(const char *jsonText, const float *boundaries, const uint64_t *seeds,
      uint32_t numExperiments, twisty_host_result *out, double *deviceMs, char *errorOut, uint32_t errorLen);

/* Gradients of one experiment's log-estimator E = log10(sum_p w_p / P). */
typedef struct twisty_host_gradients {
    double log10_estimate;  /* E, without the path normaliser (1 for N >= 7) */
    double d_eps;           /* dE / d eps, the config's eps slot (maxBound = 10/eps moves with it) */
    double d_theta, d_phi;  /* dE / d angles of the end direction (sin t cos p, sin t sin p, cos t) */
    double kernel_ms;       /* forward + adjoint device time */
    /* standard deviation of d_theta, d_phi added by the roulette of the
     * reverse sweep (twisty_hip.adjointRoulette, default 1; 0 = exact sweep,
     * and these are 0) and their Monte Carlo standard error */
    double d_theta_roulette_std, d_theta_mc_std, d_phi_roulette_std, d_phi_mc_std;
    uint64_t paths_swept;   /* paths the reverse sweep visited */
} twisty_host_gradients;

/* The adjoint of the experiment of a JSON config (text; experiment.
 * basicExperiment, experimentParams with twisty_hip.perturbationsPerPath as K,
 * paths restarting from the initial curve): dE/dT into dTableOut (nullable,
 * numCurvatureSteps doubles) and, chained on the host in FP64 through dT/d eps
 * and the initial arc's Jacobian, dE/d eps and dE/d(theta, phi).  No
 * counterpart in the reference, which has no gradients; it differentiates the
 * estimate RunnerSpecificRunExperiment returns (ExperimentRunner.h:99-101). */
int twisty_host_experiment_gradients(const char *jsonText, twisty_host_gradients *out, double *dTableOut,
      uint32_t dTableLen, char *errorOut, uint32_t errorLen);

/* ExperimentBase::Result (Experiments/ExperimentBase.h) with the weight as log10. */
typedef struct twisty_host_generation_result {
    double log10_total_weight;   /* log10 of sum_valid 10^(w + normaliser) / num_valid */
    double min_log10_w, max_log10_w;  /* with the normaliser */
    uint64_t num_valid, num_total;
    double kernel_ms;
} twisty_host_generation_result;

/* ExperimentBase::MSegmentPathGenerationMC (Experiments/ExperimentBase.cpp:
 * 1246-1386) on a JSON config (text): experimentParams give N, paths,
 * weighting, device; experiment.basicExperiment the boundary and arclength.
 * Paths are built by recursive sphere intersection (twisty_hip_run_generated;
 * N = 2 + 2^k * {2, 3}) and weighted with the table of ds = arclength / N.
 * pathNormalizerLog10 is added to every path's weight, as the reference's
 * argument of the same name (its solver passes 0). */
int twisty_host_path_generation_mc(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      twisty_host_generation_result *out, char *errorOut, uint32_t errorLen);

/* The multi-arclength overload (Experiments/ExperimentBase.cpp:1388-1597):
 * cached tables for numDsSteps ds values on [minDsTables, maxDsTables] (the
 * .cwt cache of twisty_host_cached_rt_tables under cacheRoot), ds drawn per
 * path on [minDs, maxDs] with minDs the quadratic minimum arclength over N
 * (:1437-1462); num_total = 0 when maxDs < minDs, as the reference returns. */
int twisty_host_path_generation_mc_multi(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      float minDsTables, float maxDsTables, uint32_t numDsSteps, float maxDs, const char *cacheRoot,
      twisty_host_generation_result *out, char *errorOut, uint32_t errorLen);

/* The two-table overload (Experiments/ExperimentBase.cpp:1599-1818), the
 * NoisyCircle experiments' estimator: Metropolis-Hastings chains of
 * chainLength proposals (ceil(paths / chainLength) chains; num_total = chains
 * x chainLength), ds on [minDs, minDs + maxDs] with minDs the boundary's
 * CalculateMinimumArclength over N, the cached environment tables as in
 * twisty_host_path_generation_mc_multi with absorption + scatter on every joint
 * (the reference's object tables are not read by its weighter); num_total = 0
 * when maxDs < minDs. */
int twisty_host_path_generation_mc_chains(const char *jsonText, uint64_t seed, double pathNormalizerLog10,
      float minDsTables, float maxDsTables, uint32_t numDsSteps, float maxDs, uint64_t chainLength,
      const char *cacheRoot, twisty_host_generation_result *out, char *errorOut, uint32_t errorLen);

/* The FullExperiment flow (Experiments/FullExperiment.cpp:130-266) on a JSON
 * config given as text: parse experiment.experimentParams and
 * experiment.basicExperiment, run FullExperimentRunnerOptimalPerturb_HIP.
 * On failure returns nonzero and writes a message into errorOut. */
int twisty_host_run_config(const char *jsonText, twisty_host_result *out, char *errorOut, uint32_t errorLen);

#ifdef __cplusplus
}
#endif

#endif /* TWISTY_HOST_H */
