// full_experiment_main.cpp -- the FullExperiment driver (Experiments/
// FullExperiment.cpp:130-266) on the HIP runner: reads a JSON config, runs
// experiment.basicExperiment and writes Results.dat in the same format.
#include <filesystem>
#include <fstream>
#include <iostream>

#include "twisty_host.hpp"

int main(int argc, char *argv[])
{
    if (argc < 2) {
        std::cout << "Call as: " << argv[0] << " configFilename" << std::endl;
        return 1;
    }
    try {
        const twisty::Json cfg = twisty::Json::ParseFile(argv[1]);
        auto ep = twisty::ExperimentRunner::ParseExperimentParamsFromConfig(cfg);
        if (ep.writeFiles) {
            std::filesystem::create_directories(ep.experimentDirPath);
            const std::string copy = ep.experimentDirPath + "/" + ep.experimentName + ".json";
            if (!std::filesystem::exists(copy))
                std::filesystem::copy_file(argv[1], copy, std::filesystem::copy_options::overwrite_existing);
        }
        twisty::PerturbUtils::BoundaryConditions bc = twisty::ParseBasicExperiment(cfg);
        ep.arclength = bc.arclength;
        ep.perExperimentDirSubfolder = "main";
        std::cout << "experimentDirPath: " << ep.experimentDirPath << std::endl;
        std::cout << "Experiment Path Count: " << ep.numPathsInExperiment << std::endl;
        std::cout << "Selected Runner Method: FullExperimentRunnerOptimalPerturb_HIP" << std::endl;
        twisty::FullExperimentRunnerOptimalPerturb_HIP runner(ep);
        const auto res = runner.RunExperiment(bc);
        if (!res) {
            std::cout << "Experiment failed: no results returned." << std::endl;
            return 1;
        }
        const twisty::BigWeight avg = res->experimentWeight / (double)res->totalPathsGenerated;
        std::cout << "Paths Generated: " << res->totalPathsGenerated << std::endl;
        std::cout << "\tTotal experiment weight: " << res->experimentWeight.ToString() << std::endl;
        std::cout << "\tAvg path weight: " << avg.ToString() << std::endl;
        std::cout << "\tFailed paths: " << res->numFailedPaths << std::endl;
        std::cout << "\tSetup / Perturb / Kernel (ms): " << res->setupExperimentMs << " / "
                  << res->perturbExperimentMs << " / " << res->kernelMs << std::endl;
        if (ep.writeFiles) {
            std::ofstream ofs(ep.experimentDirPath + "/Results.dat");
            ofs << "Scatter value: " << ep.weightingParameters.scatter << std::endl;
            ofs << "\tTotal experiment weight: " << res->experimentWeight.ToString() << std::endl;
            ofs << "\tAvg path weight: " << avg.ToString() << std::endl;
            ofs << "\tTotal experiment time (ms) " << res->totalExperimentMs << std::endl;
            ofs << "\tSetup time (ms) " << res->setupExperimentMs << std::endl;
            ofs << "\tPerturb time (ms) " << res->perturbExperimentMs << std::endl;
            ofs << "\tWeighting time (ms) " << res->weightingExperimentMs << std::endl;
        }
    } catch (const std::exception &e) {
        std::cout << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
