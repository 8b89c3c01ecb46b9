// barnes_hut.h -- C++ simulation object over the C ABI (the reference's
// BarnesHut class, 3D/CUDA/CUDA_NBody/include/BarnesHut.cuh:18-127).
#pragma once

#include <hip/hip_runtime.h>

#include <sstream>
#include <string>
#include <vector>

#include "../../include/nbx.h"

namespace nbx {

// Leveled log lines in the reference's format (3D/CUDA/CUDA_NBody/include/
// Logger.h:40-74, src/Logger.cpp:11-50): "[TIME ] message", printed when the
// message's level is at or above the configured one, levels ordered DEBUG <
// WARN < ERROR < INFO < TIME; the default prints TIME only.  The label is
// coloured (ANSI) when stdout is a terminal.
enum class LogLevel { DEBUG = 0, WARN = 1, ERROR = 2, INFO = 3, TIME = 4 };
struct LogConfig {
    bool headers = true;
    LogLevel level = LogLevel::TIME;
};
extern LogConfig logcfg;
// Verbosity of the reference's -v count (src/main.cu:63-78): 0 TIME, 1 INFO,
// 2 ERROR, 3 WARN, 4 or more DEBUG.
LogLevel level_of_verbosity(int count);
class Log {
public:
    explicit Log(LogLevel type);
    ~Log();
    template <class T> Log &operator<<(const T &v)
    {
        if (on_) os_ << v;
        return *this;
    }

private:
    bool on_;
    std::ostringstream os_;
};

struct SimulationParameters {  // include/Constants.h:4-22 plus the engine's knobs
    bool timeKernels = true;
    int iterations = 100;
    int numberOfParticles = 524288;
    float timestep = 0.001f;
    float gravity = 1.0f;
    float dampening = 1.0f;
    float theta = 0.5f;
    float softening = 1e-3f;
    bool kahan = true;
    bool buildLocked = false;           // subtree build by lock-based insertion
    std::string model = "binary_disk";  // star_disk | plummer | binary_disk
    unsigned long long seed = 1;
    bool copyToHost = false;            // the reference copies x, v back every step
    // The first numberOfSuns bodies of the initial conditions are kept outside
    // the tree as the reference's OpenMP code keeps its suns
    // (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:111-124); numberOfParticles
    // counts them.
    int numberOfSuns = 0;
    // The reference's launch and rendering settings (include/Constants.h:15-19).
    // blockSize sets the threads per block of the streaming stages; gridSize,
    // warp and stackSize are kept for the interface only (grids are sized from
    // N, a walk is one 64-lane wave with its stack in LDS), and nothing is
    // rendered (renderInterval, render).
    int gridSize = 1024;
    int blockSize = 256;
    int warp = 32;
    int stackSize = 64;
    int renderInterval = 10;
    bool render = false;
};

class BarnesHut {
public:
    explicit BarnesHut(const SimulationParameters &p);
    ~BarnesHut();
    BarnesHut(const BarnesHut &) = delete;
    BarnesHut &operator=(const BarnesHut &) = delete;

    // One step (src/BarnesHut.cu:195-301); throws std::runtime_error on an engine error.
    void update(int step);
    float getSystemSize() const;

    SimulationParameters parameters;
    int numParticles, numNodes;  // bodies in the tree (suns excluded) and nodes
    int numSuns;
    std::vector<float> h_x, h_y, h_z, h_vx, h_vy, h_vz, h_mass;
    std::vector<float> h_sun;  // 7 floats per sun: x, y, z, vx, vy, vz, mass
    // per-step stage times in ms: reset, bbox, build, com, sort, forces, update, copy, all
    std::vector<float> time_resetArrays, time_computeBoundingBox, time_buildTree, time_centreOfMass, time_sort,
        time_computeForces, time_update, time_copyDeviceToHost, time_all;

private:
    nbx_context *ctx = nullptr;
    hipStream_t stream = nullptr;
    float *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_mass = nullptr, *d_vx = nullptr, *d_vy = nullptr,
          *d_vz = nullptr, *d_ax = nullptr, *d_ay = nullptr, *d_az = nullptr;
    int *d_child = nullptr, *d_start = nullptr, *d_sorted = nullptr, *d_count = nullptr, *d_index = nullptr,
        *d_mutex = nullptr;
    float *d_min_x = nullptr, *d_max_x = nullptr, *d_min_y = nullptr, *d_max_y = nullptr, *d_min_z = nullptr,
          *d_max_z = nullptr;
    float *d_sun = nullptr;  // 10 arrays of numSuns: x, y, z, vx, vy, vz, mass, ax, ay, az
    void check(int rc, const char *what);
};

}  // namespace nbx
