/*
 * twisty_hip.h -- C ABI of libtwisty_hip.so, the gfx950 implementation of
 * Twisty's optimal-perturb path-integral hot path.
 *
 * Conventions: every function returns 0 (TWISTY_HIP_OK) or an error code; the
 * message of the last error of a context is twisty_hip_last_error(ctx).  No
 * exceptions cross this boundary.  All array arguments are HOST pointers and
 * every call is synchronous.  One host thread per context; contexts are
 * independent and each owns its device memory and HIP stream.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the Twisty repository root).
 */
#ifndef TWISTY_HIP_H
#define TWISTY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWISTY_HIP_OK 0
#define TWISTY_HIP_ERR_ARG 1         /* invalid argument (nothing was launched) */
#define TWISTY_HIP_ERR_HIP 2         /* a HIP runtime call failed */
#define TWISTY_HIP_ERR_STATE 3       /* call order: tables / curve not set */
#define TWISTY_HIP_ERR_UNSUPPORTED 4 /* shape outside what the kernels support */

typedef struct twisty_hip_ctx twisty_hip_ctx;

/* Result of one run.  Replaces the per-1e6-path CombinedWeightValues_C array
 * and its Boost big-float extraction (CombinedWeightUtils.h:28-174,
 * FullExperimentRunnerOptimalPerturb.cpp:342-378): the estimator numerator is
 * returned as a base-10 logarithm computed in FP64 from per-block partials. */
typedef struct twisty_hip_partials {
    double log10_sum_w;   /* log10 of sum over valid paths of 10^{w_p}; -inf if none */
    double log10_sum_w2;  /* log10 of sum over valid paths of 10^{2 w_p} */
    double max_log10_w;   /* max over valid paths of w_p */
    uint64_t num_paths;   /* paths generated (valid + invalid) */
    uint64_t num_valid;
    uint64_t num_invalid; /* a segment weight was NaN: PathWeighters.cpp:89-102 */
    uint64_t num_clamped; /* joints clamped into the table range: PathWeighters.cpp:53-65 */
} twisty_hip_partials;

/* Library version (major*10000 + minor*100 + patch). */
int twisty_hip_version(void);

/* Number of visible HIP devices.  Replaces SetupCudaDevice's device query
 * (FullExperimentRunnerOptimalPerturbOptimized_GPU.cu:395-457). */
int twisty_hip_device_count(int *count);

/* Creates a context on `device`: stream, events and the per-block partials
 * buffer.  Replaces SetupCudaDevice + the cudaMalloc half of SetupCudaPerturb
 * (FullExperimentRunnerOptimalPerturbOptimized_GPU.cu:395-457, 527-658). */
int twisty_hip_create(int device, twisty_hip_ctx **out);

/* Frees everything the context owns.  Replaces CleanupCudaPerturb /
 * CleanupCudaRandStates / CleanupCudaDevice (...Optimized_GPU.cu:370-376,
 * 660-676). */
int twisty_hip_destroy(twisty_hip_ctx *ctx);

/* Message of the last failed call on ctx ("" when none).  Never NULL. */
const char *twisty_hip_last_error(const twisty_hip_ctx *ctx);

/* Uploads the curvature -> segment-weight table(s).  The argument list is the
 * table half of WeightCurveViaCurvatureLog10_CudaSafe (PathWeighters.h:25-28,
 * PathWeighters.cpp:27-104) and of the kernel arguments
 * (...Optimized_GPU.cu:96-102), i.e. BaseWeightLookupTable's
 * AccessLookupTable / GetMinCurvature / GetMaxCurvature /
 * GetCurvatureStepSize / GetDs (PathWeightUtils.h:129-139) plus the
 * absorption coefficient.  `tables` holds numTables tables of numBins floats
 * each, sharing one curvature grid; numTables > 1 combines them by summing
 * the interpolated segment weights.  The host buffer is copied; the caller
 * keeps it.  Tables must not contain +-inf. */
int twisty_hip_set_tables(twisty_hip_ctx *ctx, const float *tables, uint32_t numBins, uint32_t numTables,
      float minCurvature, float maxCurvature, float curvatureStepSize, float ds, float absorption);

/* Uploads the initial curve: (numSegments+1) points as x,y,z floats (the
 * reference's Farlor::Vector3 positions, Curve.h:155-157) and the boundary
 * conditions in BoundaryConditions_CudaSafe layout (CurvePerturbUtils.h:70-76):
 * startPos[3], startDir[3], endPos[3], endDir[3], arclength.  Replaces the
 * per-thread curve replication of SetupCudaPerturb (...Optimized_GPU.cu:610-644)
 * with one copy; the segment vectors p_{i+1} - p_i are formed on the host in
 * FP64 and stored as floats. */
int twisty_hip_set_curve(twisty_hip_ctx *ctx, const float *positions, uint32_t numSegments,
      const float boundary[13]);

/* Runs paths [pathOffset, pathOffset + numPaths).  Each path applies
 * perturbationsPerPath endpoint-preserving sub-chain rotations, then weights
 * its N-1 joints (the loop body of FullExperimentRunnerOptimalPerturb.cpp:469-593
 * and ...Optimized_GPU.cu:735-888).  chainLength = 1 starts every path from the
 * initial curve; chainLength = L > 1 carries the state through L consecutive
 * path ids as the reference's per-thread Markov chain does: chains start at the
 * path ids that are multiples of L (pathOffset must be one), and the last chain
 * of a range may be shorter.  Draws: Philox4x32-10, key = seed,
 * counter = (perturbation, 0, path).
 * perPathLog10 (numPaths floats) and perJointCurvature (numPaths*(N-1)
 * floats) are optional host outputs (NULL to skip); an invalid path reports
 * NaN.  Replaces the kernel launch, cudaDeviceSynchronize and the D2H copy of
 * the accumulators (...Optimized_GPU.cu:268-340). */
int twisty_hip_run(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint32_t perturbationsPerPath,
      uint32_t chainLength, uint64_t seed, twisty_hip_partials *out, float *perPathLog10,
      float *perJointCurvature);

/* twisty_hip_run without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same draws, kernels,
 * schedule and result as twisty_hip_run with perPathLog10 and
 * perJointCurvature NULL: the per-block partials and the FP64 result live in a
 * static device array of the library (about 128 KB per GPU, held until the
 * process ends) instead of buffers the context allocates.  Limits: at most
 * 4096 blocks per launch (numCUs * blocksPerCU; TWISTY_HIP_ERR_UNSUPPORTED
 * beyond), the exit-direction and weight histograms and the kept final state
 * must be off (TWISTY_HIP_ERR_UNSUPPORTED otherwise).  Tables and curve come
 * from set_tables / set_curve as for twisty_hip_run.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two contexts on two host
 * threads) race and must be serialised by the caller. */
int twisty_hip_run_noalloc(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset,
      uint32_t perturbationsPerPath, uint32_t chainLength, uint64_t seed, twisty_hip_partials *out);

/* Launch schedule of twisty_hip_run: lanesPerPath in {1,2,4,8,16,32,64},
 * blockSize one of the compiled sizes for that lanesPerPath, blocksPerCU >= 1;
 * 0 for any value selects the built-in choice for the current curve.  For
 * lanesPerPath 4, blockSize 768 is the lean kernel (table read from global
 * memory, draw records exchanged between lanes without LDS), the built-in
 * choice whenever it keeps more lanes per CU resident; for lanesPerPath 32,
 * blockSize 352 keeps eleven paths per CU (the last wave half full) and 384
 * reads the table from global memory where the table-in-LDS instance does not
 * fit (N = 1024).  Replaces the
 * compile-time PerturbGridSize / PerturbBlockSize (...Optimized_GPU.cu:45-47). */
int twisty_hip_set_schedule(twisty_hip_ctx *ctx, uint32_t lanesPerPath, uint32_t blockSize,
      uint32_t blocksPerCU);

/* The schedule the next run will use for the current curve (after resolving
 * zeros) and the dynamic LDS bytes per block it needs. */
int twisty_hip_get_schedule(twisty_hip_ctx *ctx, uint32_t *lanesPerPath, uint32_t *blockSize,
      uint32_t *blocksPerCU, uint32_t *ldsBytes);

/* Device time of the last run measured with HIP events on the context's
 * stream: the forward kernel alone, and the whole call's device work (forward,
 * FP64 finalize and, when enabled, the exit histogram or the adjoint). */
int twisty_hip_last_timing(twisty_hip_ctx *ctx, float *forwardMs, float *totalMs);

/* Russian roulette in front of the reverse sweep of the following
 * twisty_hip_run_adjoint calls (the dE/d(initial segments) -> exit-direction
 * gradient): kappa > 0 keeps path p with probability q_p = min(1, omega_p *
 * numPaths / kappa) (omega_p = w_p / sum w from the forward pass; draw:
 * Philox4x32-10, counter (0xA0000000, 0, path lo, path hi), key = seed) and
 * scales its contribution by 1 / q_p, so the gradient keeps its expectation
 * while about numPaths / kappa + (paths with omega_p >= kappa / numPaths)
 * paths are swept.  The added variance is at most kappa / numPaths times the
 * omega-weighted mean square of the per-path gradients, and is reported per
 * direction by twisty_hip_last_adjoint_direction_stats next to the gradient's
 * Monte Carlo error.  The default is kappa = 1; kappa = 0 sweeps every path
 * and gives the exact gradient of the paths' estimate (the one to compare with
 * a same-draw reference).  The table gradient always takes every path.  No
 * counterpart in the reference, which has no adjoint. */
int twisty_hip_set_adjoint_roulette(twisty_hip_ctx *ctx, float kappa);

/* Paths the last twisty_hip_run_adjoint visited (numPaths without roulette). */
int twisty_hip_last_adjoint_paths(twisty_hip_ctx *ctx, uint64_t *pathsVisited);

/* Independent-path estimator: ExperimentBase::MSegmentPathGenerationMC
 * (Experiments/ExperimentBase.cpp:1246-1386).  Each path is built from scratch
 * between the hinge points of the curve's boundary (set_curve) by recursive
 * sphere intersection (ResolveEvenNumberOfSegments / ResolveThreeSegments /
 * ResolveTwoSegments, :617-983), with ds = arclength / N, and weighted with
 * the context's table as twisty_hip_run weighs a path.  N must be
 * 2 + 2^k * 2 or 2 + 2^k * 3 (k >= 1: 6, 8, 10, 14, 18, 26, 34, 50, 66, ...),
 * as the reference's recursion requires.  Draws: one Philox4x32-10 block per
 * recursion node, counter (node, 0x80000000 | sub, path lo, path hi), key =
 * seed.  A step the reference aborts with an exception (spheres that do not
 * meet) makes the path invalid.  out: the partials over paths
 * [pathOffset, pathOffset + numPaths) (the reference's estimate is
 * sum / num_valid); minLog10W (nullable): the smallest valid log10 weight;
 * perPathLog10 (nullable): every path's log10 weight, NaN when invalid;
 * perJointCurvature (nullable): numPaths * (N - 1) joint curvatures, path-major
 * (check output, as in twisty_hip_run). */
int twisty_hip_run_generated(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint64_t seed,
      twisty_hip_partials *out, double *minLog10W, float *perPathLog10, float *perJointCurvature);

/* Multi-ds tables for twisty_hip_run_generated_multi: numTables tables of
 * numBins floats, tableMeta 4 floats per table (ds, minCurvature,
 * maxCurvature, curvatureStepSize), ds strictly increasing on a uniform grid
 * (CachedMultiArclengthWeightLookupTable, PathWeightUtils.cpp:123-278), the
 * absorption coefficient and the segment count the tables will weigh. */
int twisty_hip_set_multi_tables(twisty_hip_ctx *ctx, const float *tables, uint32_t numBins, uint32_t numTables,
      const float *tableMeta, float absorption, uint32_t numSegments);

/* The multi-arclength form of the independent-path estimator
 * (Experiments/ExperimentBase.cpp:1388-1597): every path draws its own
 * ds = u (maxDs - minDs) + minDs (u from node 0's Philox words), is built with
 * that ds and weighted with the cached table nearest to it (first on ties,
 * PathWeightUtils.cpp:223-236), its curvature taken with arclength ds N and its
 * absorption with the table's ds.  Otherwise as twisty_hip_run_generated (the
 * boundary comes from set_curve; set_tables is not needed). */
int twisty_hip_run_generated_multi(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset, uint64_t seed,
      float minDs, float maxDs, twisty_hip_partials *out, double *minLog10W, float *perPathLog10,
      float *perJointCurvature);

/* The two-table Metropolis-Hastings form of the independent-path estimator
 * (Experiments/ExperimentBase.cpp:1599-1818, the NoisyCircle experiments'
 * MSegmentPathGenerationMC with environment and object tables), with chains in
 * place of the reference's OpenMP threads: chains [chainOffset, chainOffset +
 * numChains), each of chainLength proposals; proposal s of chain c draws as path
 * c * chainLength + s.  A proposal draws ds = minDs + x, x exponential of rate
 * extinction truncated to [0, maxDs] (minDs = the boundary's minimum arclength
 * / N and maxDs the reference's actualMaxDs; the reference rejects draws past
 * maxDs + minDs), is built with that ds and weighted with the cached table
 * nearest to it and the loss exp(-extinction * table ds) on every joint
 * (WeightCurveViaPositionLog10_PositionDependent, PathWeighters.cpp:105-177,
 * whose object-table branch is commented out), plus log10 of the sampling
 * weight 1 / (t exp(-t (ds - minDs) N)).  extinction = absorption + scatter
 * (float), and set_multi_tables must have been given it as its absorption.  An
 * invalid proposal is skipped; a chain takes its first valid proposal and a
 * later one when u <= 10^(w + norm) / 10^(w' + norm), and every valid proposal
 * adds the chain's current weight to the partials (so the estimate is their
 * sum over out->num_valid, with pathNormalizerLog10 added by the caller).
 * perPathLog10 (nullable): each proposal's log10 weight, sampling weight
 * included, normaliser not (NaN when invalid); chainLog10 (nullable): the
 * chain's log10 weight after each proposal (NaN before its first valid one);
 * both numChains * chainLength floats, chain-major; minLog10W: the smallest
 * chain weight summed.  Draws are build-defined (Philox; the ds draw through
 * a log formed from basic FP64 operations), not the reference's mt19937_64. */
int twisty_hip_run_generated_chains(twisty_hip_ctx *ctx, uint64_t numChains, uint64_t chainLength,
      uint64_t chainOffset, uint64_t seed, float minDs, float maxDs, float extinction, double pathNormalizerLog10,
      twisty_hip_partials *out, double *minLog10W, float *perPathLog10, float *chainLog10, float *perJointCurvature);

/* Burn-in of the following twisty_hip_run calls: every chain first applies
 * pathsPerChain unweighted paths of K perturbations (draws keyed by path id
 * 2^63 | (chain * pathsPerChain + b), chain = absolute path id / chainLength),
 * then its chainLength weighted paths.  Replaces numPathsToSkip, which the
 * reference applies at the start of every thread's chain
 * (FullExperimentRunnerOptimalPerturb.cpp:286, 469-478, 559-561).
 * 0 (the default) switches it off; twisty_hip_run_adjoint never burns in. */
int twisty_hip_set_burn_in(twisty_hip_ctx *ctx, uint32_t pathsPerChain);

/* Keep every path's final state in device memory on the following runs
 * (12*N bytes per path), for twisty_hip_get_final_positions. */
int twisty_hip_set_keep_final_state(twisty_hip_ctx *ctx, int enable);

/* Final points of the last run's paths, (numSegments+1) x,y,z floats per path,
 * path after path (the first point of the initial curve plus the running sum of
 * the final segments, accumulated in FP64).  Feeds the path-batch export of
 * FullExperimentRunnerOptimalPerturb.cpp:599-858 (.pbd).  numPaths must equal
 * the last run's path count; twisty_hip_run_adjoint keeps the final state too. */
int twisty_hip_get_final_positions(twisty_hip_ctx *ctx, float *positions, uint64_t numPaths);

/* Weight-magnitude histogram of the following runs: every valid path counted
 * in one of numBins linear bins of its weight 10^(log10 w + pathNormalizerLog10)
 * on [minWeight, maxWeight], below -> bin 0, above -> last bin (the histogram
 * FullExperimentRunnerOptimalPerturb.cpp:26-37, 235-246, 394-422, 562-575
 * keeps with 500 bins on [4.69653e-16, 1.17517e-13]).  0 bins switches it off. */
int twisty_hip_set_weight_histogram(twisty_hip_ctx *ctx, uint32_t numBins, double minWeight, double maxWeight,
      double pathNormalizerLog10);

/* Path counts of the last run's weight histogram (numBins uint64). */
int twisty_hip_get_weight_histogram(twisty_hip_ctx *ctx, uint64_t *counts);

/* Exit-direction histogram of the following runs: nTheta x nPhi bins of the
 * last free tangent t_{N-2} in (cos theta, phi) about the end direction of the
 * boundary conditions (end frame: f1 = the coordinate axis least aligned with
 * e, made orthogonal to e; f2 = e x f1; phi = atan2(t.f2, t.f1)).  Bin
 * (it, ip) is it*nPhi + ip.  Each bin holds the share of the estimator's
 * weight sum carried by its paths (FP32 compensated bin sums, FP64
 * normaliser), so the bins add up to 1.  0, 0 switches it off.
 * Build-defined: the reference's only histogram bins weight magnitudes
 * (FullExperimentRunnerOptimalPerturb.cpp:26-37, 235-246, 394-422). */
int twisty_hip_set_exit_histogram(twisty_hip_ctx *ctx, uint32_t nTheta, uint32_t nPhi);

/* Copies the histogram of the last run (nTheta*nPhi doubles) and, optionally,
 * every path's bin index (numPaths must equal the last run's path count). */
int twisty_hip_get_exit_histogram(twisty_hip_ctx *ctx, double *histogram, uint32_t *perPathBin,
      uint64_t numPaths);

/* Forward pass plus reverse-mode derivative of the log-estimator
 * E = log10(sum_p w_p / numPaths) (paths restart from the initial curve,
 * chainLength 1).  dLog10EstimateDTable (numBins doubles) receives dE/dT[b]
 * for the combined table; with several tables set, it is also the derivative
 * for each of them.  dLog10EstimateDSegments (3*N doubles, NULL to skip the
 * reverse sweep) receives dE/dd_i for the initial segment vectors
 * d_i = p_{i+1} - p_i; the host maps it onto boundary parameters (exit
 * direction).  The curvature derivative of the segment weight uses the
 * bin-average slope (T[r]-T[l])/step of the reference's lerp
 * (PathWeighters.cpp:67-82).  The reference has no derivative; the call it
 * extends is RunnerSpecificRunExperiment (ExperimentRunner.h:99-101). */
int twisty_hip_run_adjoint(twisty_hip_ctx *ctx, uint64_t numPaths, uint64_t pathOffset,
      uint32_t perturbationsPerPath, uint64_t seed, twisty_hip_partials *out, double *dLog10EstimateDTable,
      double *dLog10EstimateDSegments);

/* Directions of the initial segments along which the following
 * twisty_hip_run_adjoint calls with a reverse sweep report per-result
 * statistics: numDirections (<= 4) vectors of 3*N floats (N of the current
 * curve), e.g. d(initial segments)/d(exit angle).  0 switches it off.  No
 * counterpart in the reference. */
int twisty_hip_set_adjoint_directions(twisty_hip_ctx *ctx, const float *directions, uint32_t numDirections);

/* For the last twisty_hip_run_adjoint with directions, per direction v:
 * gradient[v] = dE/dv (dLog10EstimateDSegments projected on v),
 * rouletteStd[v] the standard deviation the roulette adds (0 for the exact
 * sweep), monteCarloStd[v] the Monte Carlo standard error of dE/dv over
 * independent path samples (delta method for the self-normalised estimator,
 * estimated from the swept paths).  Arrays of numDirections doubles, each
 * nullable. */
int twisty_hip_last_adjoint_direction_stats(twisty_hip_ctx *ctx, double *gradient, double *rouletteStd,
      double *monteCarloStd);

/* Launch schedule of the adjoint kernel (same meaning as
 * twisty_hip_set_schedule; 0 = built-in choice).  It applies to the pass
 * with the reverse sweep (to the table-only pass when there is no sweep); the
 * table pass in front of a roulette sweep takes the built-in choice. */
int twisty_hip_set_adjoint_schedule(twisty_hip_ctx *ctx, uint32_t lanesPerPath, uint32_t blockSize,
      uint32_t blocksPerCU);

/* Device time of the adjoint part (kernel + gradient reductions) of the last
 * twisty_hip_run_adjoint; 0 after a plain run. */
int twisty_hip_last_adjoint_timing(twisty_hip_ctx *ctx, float *adjointMs);

/* ------------------------------------------------------------------------
 * Batched experiments.  One persistent launch (captured once in a HIP graph and
 * replayed) runs many experiments with different segment counts, path counts,
 * geometries and tables.  Replaces a driver's serial loop of one
 * ExperimentRunner per experiment (Experiments/BeamSpreadExperiment.cpp:209-284,
 * Experiments/FullExperiment.cpp:204-232).
 * ------------------------------------------------------------------------ */
typedef struct twisty_hip_batch twisty_hip_batch;

/* One experiment of a batch: draws, chains and burn-in as twisty_hip_run with
 * path ids 0..numPaths-1, pathOffset 0, the same chainLength and burn-in. */
typedef struct twisty_hip_experiment {
    uint64_t num_paths;
    uint64_t seed;
    uint32_t num_segments;            /* N: 4 <= N <= 64 * segments_per_lane */
    uint32_t perturbations_per_path;  /* K >= 1 */
    uint32_t table_index;             /* < numTables of the batch */
    uint32_t positions_offset;        /* first point (not float) of its curve in `positions` */
    float arclength;                  /* the boundary's arclength (ds = arclength / N) */
    uint32_t chain_length;            /* paths per chain (0 or 1: every path restarts) */
    uint32_t burn_in;                 /* unweighted paths opening every chain (<= 2^24) */
} twisty_hip_experiment;

/* Launch schedule of a batch: blocksPerCU 0 = as many as LDS and registers
 * allow; segmentsPerLane picks each experiment's lanes per path (the smallest
 * power of two >= 2 with ceil(N / lanes) <= segmentsPerLane); tilePaths = paths
 * per work item; lean places the tables and the draw records: 0 both in LDS
 * (blockSize 256 or 512), 1 tables read from global memory and draw records
 * exchanged without LDS (blockSize 256 or 512), 2 tables read from global
 * memory and draw records in LDS (blockSize 256).  Every placement
 * gives the same bits per path. */
typedef struct twisty_hip_batch_schedule {
    uint32_t block_size;
    uint32_t blocks_per_cu;
    uint32_t segments_per_lane;
    uint32_t tile_paths;
    uint32_t lean;
} twisty_hip_batch_schedule;

/* Uploads a batch: experiments, numTables tables of numBins floats with their
 * grids (tableMeta: minCurvature, maxCurvature, curvatureStepSize, ds,
 * absorption per table), and the concatenated initial curves (x,y,z floats;
 * experiment e uses points positions_offset .. positions_offset + N).  Tables
 * are shared: an experiment weighs its joints with its table's curvature grid
 * and its own ds = arclength / N for curvature and absorption.  Copies
 * everything; the schedule starts at the built-in choice. */
int twisty_hip_batch_create(twisty_hip_ctx *ctx, const twisty_hip_experiment *experiments, uint32_t numExperiments,
      const float *tables, uint32_t numBins, uint32_t numTables, const float *tableMeta, const float *positions,
      uint64_t numPoints, twisty_hip_batch **out);

/* Frees the batch.  Destroying its context first releases the batch's device
 * memory; the batch then only accepts twisty_hip_batch_destroy (other calls
 * return TWISTY_HIP_ERR_STATE). */
int twisty_hip_batch_destroy(twisty_hip_batch *batch);

int twisty_hip_batch_set_schedule(twisty_hip_batch *batch, const twisty_hip_batch_schedule *schedule);
int twisty_hip_batch_get_schedule(twisty_hip_batch *batch, twisty_hip_batch_schedule *schedule);

/* Runs the batch (graph replay).  out: numExperiments partials.
 * perPathLog10 (nullable): every experiment's paths back to back, in
 * experiment order.  deviceMs (nullable): HIP-event time of the replay. */
int twisty_hip_batch_run(twisty_hip_batch *batch, twisty_hip_partials *out, float *perPathLog10, float *deviceMs);

/* Times every candidate schedule (reps replays each, best kept), leaves the
 * fastest set and returns it.  tried / triedMs (nullable, capacity cap)
 * receive the candidates and their times; *numTried their count. */
int twisty_hip_batch_autotune(twisty_hip_batch *batch, uint32_t reps, twisty_hip_batch_schedule *best,
      twisty_hip_batch_schedule *tried, float *triedMs, uint32_t cap, uint32_t *numTried);

#ifdef __cplusplus
}
#endif

#endif /* TWISTY_HIP_H */
