// nbx_runner.cpp -- command-line runner, the counterpart of the reference's
// ./bin/runner (3D/CUDA/CUDA_NBody/src/main.cu:19-190): the same options
// (-r -i -n -b -g -R -v -h, with the same defaults), per-stage [TIME] lines per
// step and the per-stage totals at the end.  -b sets the threads per block of
// the streaming stages; -g, -R and -r are accepted and have no effect (grids
// are sized from N; nothing is rendered).
//
//   nbx_runner [-n particles] [-i iterations] [-b blocksize] [-g gridsize] [-R interval] [-r] [-v]
//              [--model binary_disk|plummer|star_disk] [--theta T] [--softening E] [--seed S]
//              [--copy] [--locked-build]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "barnes_hut.h"

int main(int argc, char **argv)
{
    nbx::SimulationParameters p;
    bool verbose = false;
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
        else if (a == "-v" || a == "--verbosity") verbose = true;
        else if (a == "--model") p.model = next("--model");
        else if (a == "--theta") p.theta = (float)std::atof(next("--theta"));
        else if (a == "--softening") p.softening = (float)std::atof(next("--softening"));
        else if (a == "--seed") p.seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--copy") p.copyToHost = true;
        else if (a == "--locked-build") p.buildLocked = true;
        else if (a == "-h" || a == "--help") {
            std::printf("nbx_runner [-n particles] [-i iterations] [-b blocksize] [-g gridsize] [-R interval] [-r]\n"
                        "           [-v] [--model binary_disk|plummer|star_disk] [--theta T] [--softening E]\n"
                        "           [--seed S] [--copy] [--locked-build]\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
        }
    }
    if (p.numberOfParticles < 2 || p.iterations < 1 || p.blockSize < 64 || p.blockSize > 1024 ||
        p.blockSize % 64 != 0) {
        std::fprintf(stderr, "need -n >= 2, -i >= 1 and -b a multiple of 64 in [64, 1024]\n");
        return 1;
    }
    try {
        nbx::BarnesHut sim(p);
        for (int s = 0; s < p.iterations; ++s) {
            sim.update(s);
            if (verbose)
                std::printf("[TIME ] step %d: reset %.3f bbox %.3f build %.3f com %.3f sort %.3f forces %.3f "
                            "update %.3f all %.3f ms\n",
                            s, sim.time_resetArrays[s], sim.time_computeBoundingBox[s], sim.time_buildTree[s],
                            sim.time_centreOfMass[s], sim.time_sort[s], sim.time_computeForces[s],
                            sim.time_update[s], sim.time_all[s]);
        }
        const char *names[] = {"reset arrays", "compute bounding boxes", "build tree", "compute COM", "sort",
                               "compute forces", "update bodies", "copy from device to host", "TOTAL"};
        const std::vector<float> *series[] = {&sim.time_resetArrays, &sim.time_computeBoundingBox,
                                              &sim.time_buildTree, &sim.time_centreOfMass, &sim.time_sort,
                                              &sim.time_computeForces, &sim.time_update,
                                              &sim.time_copyDeviceToHost, &sim.time_all};
        for (int k = 0; k < 9; ++k) {
            double tot = 0;
            for (float v : *series[k]) tot += v;
            std::printf("[TIME ] Time to %s: %.3f ms\n[TIME ] \tper step: %.4f ms\n", names[k], tot,
                        tot / p.iterations);
        }
        std::printf("[INFO ] Number of particles: %d\n[INFO ] Number of iterations: %d\n", p.numberOfParticles,
                    p.iterations);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[ERROR] %s\n", e.what());
        return 2;
    }
    return 0;
}
