// beam_spread_main.cpp -- the BeamSpreadExperiment driver
// (Experiments/BeamSpreadExperiment.cpp:130-298) with every experiment of its
// loops in one BatchedExperimentRunner launch instead of one runner each.
//
// Geometry as the reference: the receiver direction from std::mt19937
// (targetSeed) and uniform floats (:183-194), the receiver on a sphere of
// sphereRadius around startPos, the minimum arclength of that boundary (:201-204).
// Seeds as the reference draws them: one mt19937(bootstrapSeed) stream of
// nonzero initial-curve seeds per arclength and one mt19937(perturbSeed)
// stream of nonzero per-curve seeds per initial curve (:212-230).  The value
// is the mean over arclengths of the mean over initial curves of the largest
// experiment weight over per-curve seeds (:276-285).
//
// Where the reference's loop does not run as written, this driver does what
// the loop states:
//  * the reference hands RunExperiment the boundary without the arclength it
//    just computed (:241-249, :268); here every experiment gets its arclength;
//  * the reference draws both seeds and passes neither; here experiment
//    (i, j) runs with seed initialCurveSeed << 32 | perCurveSeed (the
//    initial curve itself is the circular arc for every i);
//  * numArclengths is fixed to 1 at :168 and the grid min + step*idx would put
//    its one experiment at the minimum arclength, where the only curve is the
//    straight/arc limit; here numArclengths is read from the config (default
//    1) and the arclengths are the midpoints min + step*(idx + 1/2);
//  * beamSpreadValues.dat, which the reference opens and leaves empty
//    (:288-295), receives the value and one line per experiment.
// --plan prints the experiments without running them.
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <random>
#include <string>

#include "twisty_host.hpp"

namespace {
constexpr float kPi = 3.141592653589793f;  // MathConsts.h:8

void process_random_seed(uint32_t &seed)  // ExperimentRunner.cpp:40-46
{
    if (seed == 0) seed = (uint32_t)std::time(nullptr);
}

uint32_t next_nonzero(std::mt19937 &g)
{
    uint32_t s = 0;
    do {
        s = g();
    } while (s == 0);
    return s;
}

double log10_add(double a, double b)
{
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    const double m = std::max(a, b);
    return m + std::log10(std::pow(10.0, a - m) + std::pow(10.0, b - m));
}
}  // namespace

int main(int argc, char *argv[])
{
    if (argc < 2) {
        std::cout << "Call as: " << argv[0] << " configFilename [--plan]" << std::endl;
        return 1;
    }
    const bool planOnly = argc > 2 && std::string(argv[2]) == "--plan";
    try {
        const twisty::Json cfg = twisty::Json::ParseFile(argv[1]);
        auto ep = twisty::ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        process_random_seed(ep.bootstrapSeed);
        process_random_seed(ep.curvePurturbSeed);
        const twisty::Json &b = cfg.at("experiment").at("beamSpreadExperiment");
        auto v3 = [&](const char *k) {
            const twisty::Json &a = b.at(k);
            return twisty::Vec3((float)a.at(0).number(), (float)a.at(1).number(), (float)a.at(2).number());
        };
        const float sphereRadius = (float)b.at("sphereRadius").number();
        const float maxArclength = (float)b.at("maxArclength").number();
        const twisty::Vec3 startPos = v3("startPos");
        const twisty::Vec3 startDir = v3("startDir");
        const uint32_t numArclengths = b.contains("numArclengths") ? (uint32_t)b.at("numArclengths").number() : 1u;
        const uint32_t numInitialCurves = (uint32_t)b.at("numInitialCurves").number();
        const uint32_t numPerInitialCurve = (uint32_t)b.at("numPerInitialCurve").number();
        uint32_t targetSeed = (uint32_t)b.at("targetSeed").number();
        if (numArclengths == 0 || numInitialCurves == 0 || numPerInitialCurve == 0)
            throw std::runtime_error("numArclengths, numInitialCurves and numPerInitialCurve must be >= 1");
        process_random_seed(targetSeed);
        std::mt19937 randomGenerator(targetSeed);
        std::uniform_real_distribution<float> uniformFloatGen(0.0f, 1.0f);
        const float n1 = (uniformFloatGen(randomGenerator) * 2.0f) - 1.0f;
        const float e1 = uniformFloatGen(randomGenerator);
        const float rxy = std::sqrt(1.0f - n1 * n1);
        const twisty::Vec3 endDir(n1, std::cos(2.0f * kPi * e1) * rxy, std::sin(2.0f * kPi * e1) * rxy);
        const twisty::Vec3 endPos = startPos + endDir * sphereRadius;

        twisty::PerturbUtils::BoundaryConditions bc;
        bc.m_startPos = startPos;
        bc.m_startDir = startDir.Normalized();
        bc.m_endPos = endPos;
        bc.m_endDir = endDir.Normalized();
        const uint32_t N = ep.numSegmentsPerCurve;
        const float minArclength = twisty::PathGeneration::CalculateMinimumArclength(bc, N);
        const float step = (maxArclength - minArclength) / (float)numArclengths;
        std::cout << "Min arclength: " << minArclength << std::endl;

        struct Slot {
            uint32_t a, i, j;
        };
        std::vector<twisty::BatchedExperimentRunner::Experiment> exps;
        std::vector<Slot> slots;
        for (uint32_t a = 0; a < numArclengths; ++a) {
            const float arclength = minArclength + step * ((float)a + 0.5f);
            std::mt19937 initialCurveGen(ep.bootstrapSeed);
            for (uint32_t i = 0; i < numInitialCurves; ++i) {
                const uint32_t initialCurveSeed = next_nonzero(initialCurveGen);
                std::mt19937 perCurveGen(ep.curvePurturbSeed);
                for (uint32_t j = 0; j < numPerInitialCurve; ++j) {
                    const uint32_t perCurveSeed = next_nonzero(perCurveGen);
                    twisty::BatchedExperimentRunner::Experiment x;
                    x.boundaryConditions = bc;
                    x.boundaryConditions.arclength = arclength;
                    x.seed = ((uint64_t)initialCurveSeed << 32) | perCurveSeed;
                    exps.push_back(x);
                    slots.push_back({ a, i, j });
                }
            }
        }
        std::cout.precision(9);
        std::cout << "endPos: " << endPos.x << " " << endPos.y << " " << endPos.z << std::endl;
        std::cout << "endDir: " << bc.m_endDir.x << " " << bc.m_endDir.y << " " << bc.m_endDir.z << std::endl;
        if (planOnly) {
            for (size_t k = 0; k < exps.size(); ++k)
                std::cout << "experiment " << slots[k].a << " " << slots[k].i << " " << slots[k].j << " arclength "
                          << exps[k].boundaryConditions.arclength << " seed " << exps[k].seed << std::endl;
            return 0;
        }
        std::cout << "Experiments: " << exps.size() << " x " << ep.numPathsInExperiment << " paths, one batch"
                  << std::endl;
        twisty::BatchedExperimentRunner runner(ep);
        const auto res = runner.RunExperiments(exps);
        if (!res) {
            std::cout << "Experiment failed: " << runner.LastError() << std::endl;
            return 1;
        }
        // mean over arclengths of the mean over initial curves of the max over seeds
        std::vector<double> maxOf((size_t)numArclengths * numInitialCurves, -INFINITY);
        for (size_t k = 0; k < exps.size(); ++k) {
            double &m = maxOf[(size_t)slots[k].a * numInitialCurves + slots[k].i];
            m = std::max(m, (*res)[k].experimentWeight.log10);
        }
        double value = -INFINITY;
        for (double m : maxOf) value = log10_add(value, m - std::log10((double)numInitialCurves * numArclengths));
        const std::filesystem::path outDir = std::filesystem::current_path() / ep.experimentName;
        std::filesystem::create_directories(outDir);
        std::ofstream out(outDir / "beamSpreadValues.dat");
        if (!out.is_open()) throw std::runtime_error("Failed to create output file: " + (outDir / "beamSpreadValues.dat").string());
        out.precision(17);
        out << "Beam spread value: " << twisty::BigWeight { value }.ToString() << "\n";
        out << "log10: " << value << "\n";
        out << "# arclengthIdx initialCurveIdx perCurveIdx arclength seed log10ExperimentWeight failedPaths\n";
        for (size_t k = 0; k < exps.size(); ++k) {
            const auto &r = (*res)[k];
            out << slots[k].a << " " << slots[k].i << " " << slots[k].j << " " << exps[k].boundaryConditions.arclength
                << " " << exps[k].seed << " " << r.experimentWeight.log10 << " " << r.numFailedPaths << "\n";
            std::cout << "experiment " << slots[k].a << " " << slots[k].i << " " << slots[k].j << ": log10 weight "
                      << r.experimentWeight.log10 << ", failed " << r.numFailedPaths << std::endl;
        }
        std::cout << "Beam spread value: " << twisty::BigWeight { value }.ToString() << std::endl;
        std::cout << "Device ms: " << runner.DeviceMs() << ", tables: " << runner.NumTables() << std::endl;
        std::cout << "Experiment done" << std::endl;
    } catch (const std::exception &e) {
        std::cout << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
