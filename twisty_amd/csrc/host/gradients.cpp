// gradients.cpp -- ComputeExperimentGradients: the adjoint of one experiment
// (twisty_hip_run_adjoint) chained on the host in FP64 to the phase width and
// the exit direction.
#include <cmath>
#include <ctime>

#include "../../../include/twisty_hip.h"
#include "twisty_host.hpp"

namespace twisty {

std::optional<ExperimentGradients> ComputeExperimentGradients(const ExperimentRunner::ExperimentParameters &ep,
      const PerturbUtils::BoundaryConditions &bc, std::string &error)
{
    error.clear();
    if (ep.weightingParameters.weightingMethod != WeightingMethod::RadiativeTransfer) {
        error = "only the radiative-transfer weighting (weightFunction 0) has a gradient";
        return std::nullopt;
    }
    const uint32_t N = ep.numSegmentsPerCurve;
    const uint64_t P = ep.numPathsInExperiment;
    if (N < 4 || P == 0 || !(bc.arclength > 0.0f)) {
        error = "need numSegments >= 4, pathsToGenerate >= 1 and a positive arclength";
        return std::nullopt;
    }
    const float ds = bc.arclength / (float)N;
    const int threads = ep.maxWeightThreads > 0 ? ep.maxWeightThreads : 0;
    PathWeighting::WeightLookupTableIntegral table(ep.weightingParameters, ds);
    std::vector<double> dTdEps;
    try {
        table.Initialize(threads);
        dTdEps = PathWeighting::WeightLookupTableIntegral::EpsDerivative(ep.weightingParameters, ds, true, threads);
    } catch (const std::exception &e) {
        error = e.what();
        return std::nullopt;
    }
    std::vector<Vec3> arc;
    std::vector<double> jTheta, jPhi;
    if (!PathGeneration::InitialArc(bc, N, arc) || !PathGeneration::InitialArcExitJacobian(bc, N, jTheta, jPhi)) {
        error = "no initial curve: arclength " + std::to_string(bc.arclength) + " is below the minimum " +
                std::to_string(PathGeneration::CalculateMinimumArclength(bc, N));
        return std::nullopt;
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
    // exit-angle directions of the initial segments: segment i = p_{i+1} - p_i
    // moves by J_{i+1} - J_i
    std::vector<float> dirs(2ull * 3 * N);
    for (uint32_t i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            dirs[3 * i + a] = (float)(jTheta[3 * (i + 1) + a] - jTheta[3 * i + a]);
            dirs[3ull * N + 3 * i + a] = (float)(jPhi[3 * (i + 1) + a] - jPhi[3 * i + a]);
        }
    if (rc == TWISTY_HIP_OK) rc = twisty_hip_set_adjoint_directions(ctx, dirs.data(), 2);
    if (rc == TWISTY_HIP_OK) rc = twisty_hip_set_adjoint_roulette(ctx, ep.adjointRoulette);
    const uint64_t seed = ep.curvePurturbSeed ? ep.curvePurturbSeed : (uint64_t)std::time(nullptr);
    ExperimentGradients g;
    g.dTable.assign(tv.size(), 0.0);
    std::vector<double> dSeg(3ull * N, 0.0);
    twisty_hip_partials part {};
    if (rc == TWISTY_HIP_OK)
        rc = twisty_hip_run_adjoint(ctx, P, 0, ep.perturbationsPerPath, seed, &part, g.dTable.data(), dSeg.data());
    if (rc != TWISTY_HIP_OK) {
        error = twisty_hip_last_error(ctx);
        return std::nullopt;
    }
    double dg[2], rs[2], ms[2];
    if (twisty_hip_last_adjoint_direction_stats(ctx, dg, rs, ms) == TWISTY_HIP_OK) {
        g.dThetaRouletteStd = rs[0];
        g.dThetaMcStd = ms[0];
        g.dPhiRouletteStd = rs[1];
        g.dPhiMcStd = ms[1];
    }
    twisty_hip_last_adjoint_paths(ctx, &g.pathsSwept);
    float fwdMs = 0.0f, totMs = 0.0f;
    twisty_hip_last_timing(ctx, &fwdMs, &totMs);
    g.kernelMs = totMs;
    g.log10Estimate = part.log10_sum_w - std::log10((double)P);
    for (size_t b = 0; b < tv.size(); ++b) g.dEps += g.dTable[b] * dTdEps[b];
    // segment i = p_{i+1} - p_i moves by J_{i+1} - J_i
    for (uint32_t i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            const double s = dSeg[3 * i + a];
            g.dTheta += s * (jTheta[3 * (i + 1) + a] - jTheta[3 * i + a]);
            g.dPhi += s * (jPhi[3 * (i + 1) + a] - jPhi[3 * i + a]);
        }
    return g;
}

}  // namespace twisty
