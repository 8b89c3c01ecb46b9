// nbx_runner.cpp -- command-line runner, the counterpart of the reference's
// ./bin/runner (3D/CUDA/CUDA_NBody/src/main.cu:19-190): the same options
// (-r -i -n -b -g -R -v -h, with the same defaults), the reference's log lines
// (per-stage and per-step [TIME ] lines of src/BarnesHut.cu:210-298, the
// totals of main.cu:152-181) and its verbosity levels (-v given 0..4 times,
// main.cu:63-78).  -b sets the threads per block of the streaming stages; -g,
// -R and -r are accepted and have no effect (grids are sized from N; nothing
// is rendered).
//
//   nbx_runner [-n particles] [-i iterations] [-b blocksize] [-g gridsize] [-R interval] [-r] [-v...]
//              [--model binary_disk|plummer|star_disk] [--theta T] [--softening E] [--seed S]
//              [--copy] [--locked-build] [--suns K]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "barnes_hut.h"

using nbx::Log;
using nbx::LogLevel;

int main(int argc, char **argv)
{
    nbx::SimulationParameters p;
    int verbosity = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", name);
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "-n" || a == "--particles") p.numberOfParticles = std::atoi(next("-n"));
        else if (a == "-i" || a == "--iterations") p.iterations = std::atoi(next("-i"));
        else if (a == "-b" || a == "--blocksize") p.blockSize = std::atoi(next("-b"));
        else if (a == "-g" || a == "--gridsize") p.gridSize = std::atoi(next("-g"));
        else if (a == "-R" || a == "--renderinterval") p.renderInterval = std::atoi(next("-R"));
        else if (a == "-r" || a == "--render") p.render = true;
        else if (a == "--verbosity") ++verbosity;
        else if (a.size() >= 2 && a[0] == '-' && a.find_first_not_of('v', 1) == std::string::npos)
            verbosity += (int)a.size() - 1;  // -v, -vv, -vvv: a repeat count
        else if (a == "--model") p.model = next("--model");
        else if (a == "--theta") p.theta = (float)std::atof(next("--theta"));
        else if (a == "--softening") p.softening = (float)std::atof(next("--softening"));
        else if (a == "--seed") p.seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--copy") p.copyToHost = true;
        else if (a == "--locked-build") p.buildLocked = true;
        else if (a == "--suns") p.numberOfSuns = std::atoi(next("--suns"));
        else if (a == "-h" || a == "--help") {
            std::printf("nbx_runner [-n particles] [-i iterations] [-b blocksize] [-g gridsize] [-R interval] [-r]\n"
                        "           [-v...] [--model binary_disk|plummer|star_disk] [--theta T] [--softening E]\n"
                        "           [--seed S] [--copy] [--locked-build] [--suns K]\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
        }
    }
    if (p.numberOfParticles < 2 || p.iterations < 1 || p.blockSize < 64 || p.blockSize > 1024 ||
        p.blockSize % 64 != 0 || p.numberOfSuns < 0 || p.numberOfSuns > NBX_MAX_STARS ||
        p.numberOfParticles - p.numberOfSuns < 2) {
        std::fprintf(stderr, "need -n >= 2, -i >= 1, -b a multiple of 64 in [64, 1024] and 0 <= --suns <= %d "
                             "leaving >= 2 bodies\n", NBX_MAX_STARS);
        return 1;
    }
    nbx::logcfg.headers = true;
    nbx::logcfg.level = nbx::level_of_verbosity(verbosity);
    Log(LogLevel::DEBUG) << "DEBUG output";
    Log(LogLevel::WARN) << "WARN output";
    Log(LogLevel::ERROR) << "ERROR output";
    Log(LogLevel::INFO) << "INFO output";
    Log(LogLevel::TIME) << "TIME output";
    try {
        nbx::BarnesHut sim(p);
        for (int s = 0; s < p.iterations; ++s) sim.update(s);
        float tot[9] = {};
        const std::vector<float> *series[9] = {&sim.time_resetArrays, &sim.time_computeBoundingBox,
                                               &sim.time_buildTree, &sim.time_centreOfMass, &sim.time_sort,
                                               &sim.time_computeForces, &sim.time_update,
                                               &sim.time_copyDeviceToHost, &sim.time_all};
        for (int k = 0; k < 9; ++k)
            for (float v : *series[k]) tot[k] += v;
        const float it = (float)p.iterations;
        Log(LogLevel::INFO) << "----------------------FINISHED----------------------";
        Log(LogLevel::INFO) << "";
        const char *names[8] = {"reset arrays", "compute bounding boxes", "build tree", "compute COM", "sort",
                                "compute forces", "update bodies", "copy from device to host"};
        for (int k = 0; k < 8; ++k) {
            Log(LogLevel::TIME) << "Time to " << names[k] << ": " << tot[k] << "ms";
            Log(LogLevel::TIME) << "\tper step: " << tot[k] / it << "ms";
        }
        Log(LogLevel::TIME) << "----------------------------------------------";
        Log(LogLevel::TIME) << "TOTAL TIME: " << tot[8] << "ms";
        Log(LogLevel::TIME) << "\tper step: " << tot[8] / it << "ms";
        Log(LogLevel::TIME) << "TOTAL TIME (without copying): " << tot[8] - tot[7] << "ms";
        Log(LogLevel::TIME) << "\tper step: " << (tot[8] - tot[7]) / it << "ms";
        Log(LogLevel::TIME) << "----------------------------------------------";
        Log(LogLevel::INFO) << "Number of particles: " << p.numberOfParticles;
        Log(LogLevel::INFO) << "Number of iterations: " << p.iterations;
        if (p.numberOfSuns > 0) Log(LogLevel::INFO) << "Number of suns (outside the tree): " << p.numberOfSuns;
    } catch (const std::exception &e) {
        Log(LogLevel::ERROR) << e.what();
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 2;
    }
    return 0;
}
