// ring_benchmark_main.cpp -- the RingBenchmark driver
// (Experiments/RingBenchmark.cpp:130-282) on BatchedExperimentRunner.
//
// The reference builds one receiver, the last of numZValues heights between
// zMin and zMax ("TODO: This likely should be a for loop", :186-188), at
// (ringRadius, 0, z) facing away from the emitter, with arclength
// max(1.1 |receiver - emitter|, 3) (:190-197), and runs it.  Here every
// height of that grid is one experiment of a single batched launch (the
// loop the TODO asks for); the reference's report is printed for the last
// height, and ringBenchmark.dat in the experiment directory lists every
// height.  --plan prints the experiments without running them.
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>

#include "twisty_host.hpp"

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
        const twisty::Json &r = cfg.at("experiment").at("ringBenchmark");
        auto v3 = [&](const char *k) {
            const twisty::Json &a = r.at(k);
            return twisty::Vec3((float)a.at(0).number(), (float)a.at(1).number(), (float)a.at(2).number());
        };
        twisty::PerturbUtils::BoundaryConditions geom;
        geom.m_startPos = v3("startPos");
        geom.m_startDir = v3("startDir").Normalized();
        const double zMin = r.at("zMin").number();
        const double zMax = r.at("zMax").number();
        const uint32_t numZValues = (uint32_t)r.at("numZValues").number();
        const double ringRadius = r.at("ringRadius").number();
        if (numZValues == 0) throw std::runtime_error("numZValues must be >= 1");
        const double deltaZ = numZValues > 1 ? (zMax - zMin) / (numZValues - 1) : 0.0;
        std::vector<twisty::BatchedExperimentRunner::Experiment> exps(numZValues);
        for (uint32_t zIdx = 0; zIdx < numZValues; ++zIdx) {
            const double receiverZ = zMin + deltaZ * zIdx;
            twisty::PerturbUtils::BoundaryConditions bc = geom;
            bc.m_endPos = twisty::Vec3((float)ringRadius, 0.0f, (float)receiverZ);
            bc.m_endDir = (bc.m_endPos - bc.m_startPos).Normalized();
            bc.arclength = std::max((bc.m_endPos - bc.m_startPos).Magnitude() * 1.1f, 3.0f);
            exps[zIdx].boundaryConditions = bc;
            exps[zIdx].seed = ep.curvePurturbSeed;
        }
        if (planOnly) {
            std::cout.precision(9);
            for (uint32_t z = 0; z < numZValues; ++z) {
                const auto &bc = exps[z].boundaryConditions;
                std::cout << "experiment " << z << " endPos " << bc.m_endPos.x << " " << bc.m_endPos.y << " "
                          << bc.m_endPos.z << " arclength " << bc.arclength << std::endl;
            }
            return 0;
        }
        if (ep.writeFiles) {
            std::filesystem::create_directories(ep.experimentDirPath);
            const std::string copy = ep.experimentDirPath + "/" + ep.experimentName + ".json";
            if (!std::filesystem::exists(copy))
                std::filesystem::copy_file(argv[1], copy, std::filesystem::copy_options::overwrite_existing);
        }
        std::cout << "experimentDirPath: " << ep.experimentDirPath << std::endl;
        std::cout << "Experiment Path Count: " << ep.numPathsInExperiment << std::endl;
        std::cout << "Receivers: " << numZValues << ", one batch" << std::endl;
        twisty::BatchedExperimentRunner runner(ep);
        const auto res = runner.RunExperiments(exps);
        if (!res) {
            std::cout << "Experiment failed: " << runner.LastError() << std::endl;
            return 1;
        }
        std::cout.precision(10);
        for (uint32_t z = 0; z < numZValues; ++z)
            std::cout << "z " << exps[z].boundaryConditions.m_endPos.z << ": log10 weight "
                      << (*res)[z].experimentWeight.log10 << ", failed " << (*res)[z].numFailedPaths << std::endl;
        const auto &last = res->back();
        const twisty::BigWeight avg = last.experimentWeight / (double)last.totalPathsGenerated;
        std::cout << "Paths Generated: " << last.totalPathsGenerated << std::endl;
        std::cout << "\tTotal experiment weight: " << last.experimentWeight.ToString() << std::endl;
        std::cout << "\tAvg path weight: " << avg.ToString() << std::endl;
        std::cout << "Scatter value: " << ep.weightingParameters.scatter << std::endl;
        std::cout << "\tTotal experiment time (ms) " << last.totalExperimentMs << std::endl;
        std::cout << "\tSetup time (ms) " << last.setupExperimentMs << std::endl;
        std::cout << "\tPerturb time (ms) " << last.perturbExperimentMs << std::endl;
        std::cout << "\tDevice time (ms) " << runner.DeviceMs() << std::endl;
        if (ep.writeFiles) {
            std::ofstream out(ep.experimentDirPath + "/ringBenchmark.dat");
            if (!out.is_open()) throw std::runtime_error("cannot write ringBenchmark.dat in " + ep.experimentDirPath);
            out.precision(17);
            out << "# z arclength log10ExperimentWeight log10AvgPathWeight failedPaths\n";
            for (uint32_t z = 0; z < numZValues; ++z) {
                const auto &x = (*res)[z];
                out << exps[z].boundaryConditions.m_endPos.z << " " << exps[z].boundaryConditions.arclength << " "
                    << x.experimentWeight.log10 << " " << (x.experimentWeight / (double)x.totalPathsGenerated).log10
                    << " " << x.numFailedPaths << "\n";
            }
        }
    } catch (const std::exception &e) {
        std::cout << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
