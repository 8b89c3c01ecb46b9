// barnes_hut.cpp -- the reference's BarnesHut simulation object
// (3D/CUDA/CUDA_NBody/src/BarnesHut.cu:12-301) over the nbx C ABI, with the
// reference's per-stage log lines.
#include "barnes_hut.h"

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>

namespace nbx {

LogConfig logcfg;

LogLevel level_of_verbosity(int count)
{
    switch (count) {
    case 0: return LogLevel::TIME;
    case 1: return LogLevel::INFO;
    case 2: return LogLevel::ERROR;
    case 3: return LogLevel::WARN;
    default: return LogLevel::DEBUG;
    }
}

Log::Log(LogLevel type) : on_((int)type >= (int)logcfg.level)
{
    if (!on_ || !logcfg.headers) return;
    static const char *label[] = {"[DEBUG] ", "[WARN ] ", "[ERROR] ", "[INFO ] ", "[TIME ] "};
    static const int colour[] = {90, 33, 31, 92, 34};  // dark grey, yellow, red, light green, blue
    static const bool tty = isatty(fileno(stdout)) != 0;
    if (tty) os_ << "\033[" << colour[(int)type] << "m" << label[(int)type] << "\033[39m";
    else os_ << label[(int)type];
}

Log::~Log()
{
    if (on_) std::cout << os_.str() << std::endl;
}

namespace {
template <class T> T *dmalloc(size_t count)
{
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    return static_cast<T *>(p);
}
}  // namespace

BarnesHut::BarnesHut(const SimulationParameters &p) : parameters(p)
{
    const int total = p.numberOfParticles;
    numSuns = p.numberOfSuns;
    if (numSuns < 0 || numSuns > NBX_MAX_STARS || numSuns >= total)
        throw std::runtime_error("suns: need 0 <= suns <= NBX_MAX_STARS and fewer than the bodies");
    std::vector<float> ic[7];
    for (auto &v : ic) v.assign(total, 0.f);
    int rc;
    if (p.model == "plummer")
        rc = nbx_ic_plummer(total, p.seed, ic[0].data(), ic[1].data(), ic[2].data(), ic[3].data(), ic[4].data(),
                            ic[5].data(), ic[6].data());
    else if (p.model == "star_disk")
        rc = nbx_ic_star_disk(total, p.seed, ic[0].data(), ic[1].data(), ic[2].data(), ic[3].data(), ic[4].data(),
                              ic[5].data(), ic[6].data());
    else
        rc = nbx_ic_binary_disk(total, p.seed, ic[0].data(), ic[1].data(), ic[2].data(), ic[3].data(), ic[4].data(),
                                ic[5].data(), ic[6].data());
    if (rc != NBX_OK) throw std::runtime_error("initial conditions: bad parameters");
    // the first numSuns bodies become the suns, the rest are the tree's bodies
    numParticles = total - numSuns;
    numNodes = 2 * numParticles + 12000;  // src/BarnesHut.cu:18
    const size_t n = numParticles, m = numNodes;
    std::vector<float> *host[7] = {&h_x, &h_y, &h_z, &h_vx, &h_vy, &h_vz, &h_mass};
    h_sun.assign(7 * (size_t)numSuns, 0.f);
    for (int k = 0; k < 7; ++k) {
        host[k]->assign(ic[k].begin() + numSuns, ic[k].end());
        for (int s = 0; s < numSuns; ++s) h_sun[7 * s + k] = ic[k][s];
    }
    for (auto *v : {&time_resetArrays, &time_computeBoundingBox, &time_buildTree, &time_centreOfMass, &time_sort,
                    &time_computeForces, &time_update, &time_copyDeviceToHost, &time_all})
        v->assign(p.iterations, 0.f);
    if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate failed");
    for (float **d : {&d_x, &d_y, &d_z, &d_mass, &d_vx, &d_vy, &d_vz, &d_ax, &d_ay, &d_az}) *d = dmalloc<float>(m);
    d_child = dmalloc<int>(8 * m);
    d_start = dmalloc<int>(m);
    d_sorted = dmalloc<int>(m);
    d_count = dmalloc<int>(m);
    d_index = dmalloc<int>(1);
    d_mutex = dmalloc<int>(1);
    for (float **d : {&d_min_x, &d_max_x, &d_min_y, &d_max_y, &d_min_z, &d_max_z}) *d = dmalloc<float>(1);
    for (float *d : {d_ax, d_ay, d_az}) hipMemset(d, 0, sizeof(float) * m);
    float *dev[7] = {d_x, d_y, d_z, d_vx, d_vy, d_vz, d_mass};
    for (int k = 0; k < 7; ++k) hipMemcpy(dev[k], host[k]->data(), 4 * n, hipMemcpyHostToDevice);
    if (numSuns > 0) {
        d_sun = dmalloc<float>(10 * (size_t)numSuns);
        hipMemset(d_sun, 0, sizeof(float) * 10 * numSuns);
        for (int k = 0; k < 7; ++k)
            for (int s = 0; s < numSuns; ++s)
                hipMemcpy(d_sun + k * numSuns + s, &h_sun[7 * s + k], 4, hipMemcpyHostToDevice);
    }
    nbx_params np;
    nbx_default_params(&np);
    np.theta = p.theta;
    np.softening = p.softening;
    np.gravity = p.gravity;
    np.timestep = p.timestep;
    np.dampening = p.dampening;
    np.kahan = p.kahan ? 1 : 0;
    np.build_locked = p.buildLocked ? 1 : 0;
    np.block_size = p.blockSize;
    check(nbx_create(&np, numParticles, numNodes, 0, &ctx), "nbx_create");
}

BarnesHut::~BarnesHut()
{
    if (ctx) nbx_destroy(ctx);
    for (void *d : {(void *)d_x, (void *)d_y, (void *)d_z, (void *)d_mass, (void *)d_vx, (void *)d_vy,
                    (void *)d_vz, (void *)d_ax, (void *)d_ay, (void *)d_az, (void *)d_child, (void *)d_start,
                    (void *)d_sorted, (void *)d_count, (void *)d_index, (void *)d_mutex, (void *)d_min_x,
                    (void *)d_max_x, (void *)d_min_y, (void *)d_max_y, (void *)d_min_z, (void *)d_max_z,
                    (void *)d_sun})
        if (d) hipFree(d);
    if (stream) hipStreamDestroy(stream);
}

void BarnesHut::check(int rc, const char *what)
{
    if (rc != NBX_OK) throw std::runtime_error(std::string(what) + ": " + nbx_last_error(ctx));
}

// The stage sequence of src/BarnesHut.cu:195-301 with its log lines; with suns,
// their forces follow the tree forces (timed with them) and they move with the
// bodies (timed with the update).
void BarnesHut::update(int step)
{
    const int n = numParticles, m = numNodes, ns = numSuns;
    const bool t = parameters.timeKernels;
    float ms = 0.f, ms2 = 0.f;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, stream);
    check(nbx_reset_arrays(ctx, d_mutex, d_x, d_y, d_z, d_mass, d_count, d_start, d_sorted, d_child, d_index, d_min_x,
                           d_max_x, d_min_y, d_max_y, d_min_z, d_max_z, n, m, stream, t ? &ms : nullptr),
          "resetArrays");
    time_resetArrays[step] = ms;
    if (t) Log(LogLevel::TIME) << "\tReset arrays: " << ms << " ms";
    check(nbx_compute_bounding_box(ctx, d_mutex, d_x, d_y, d_z, d_min_x, d_max_x, d_min_y, d_max_y, d_min_z,
                                   d_max_z, n, stream, t ? &ms : nullptr),
          "computeBoundingBox");
    time_computeBoundingBox[step] = ms;
    if (t) Log(LogLevel::TIME) << "\tBounding box: " << ms << " ms";
    check(nbx_build_tree(ctx, d_x, d_y, d_z, d_mass, d_count, d_start, d_child, d_index, d_min_x, d_max_x, d_min_y,
                         d_max_y, d_min_z, d_max_z, n, m, stream, t ? &ms : nullptr),
          "buildTree");
    time_buildTree[step] = ms;
    if (t) Log(LogLevel::TIME) << "\tBuilding tree: " << ms << " ms";
    check(nbx_centre_of_mass(ctx, d_x, d_y, d_z, d_mass, d_index, n, stream, t ? &ms : nullptr), "centreOfMass");
    time_centreOfMass[step] = ms;
    if (t) Log(LogLevel::TIME) << "\tCenter of mass: " << ms << " ms";
    check(nbx_sort(ctx, d_count, d_start, d_sorted, d_child, d_index, n, stream, t ? &ms : nullptr), "sort");
    time_sort[step] = ms;
    if (t) Log(LogLevel::TIME) << "\tSort particles: " << ms << " ms";
    check(nbx_compute_forces(ctx, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_ax, d_ay, d_az, d_mass, d_sorted, d_child,
                             d_min_x, d_max_x, n, parameters.gravity, stream, t ? &ms : nullptr),
          "computeForces");
    float *S = d_sun;  // x, y, z, vx, vy, vz, mass, ax, ay, az of the suns
    const size_t k = (size_t)ns;
    ms2 = 0.f;
    if (ns > 0)
        check(nbx_star_forces(ctx, S, S + k, S + 2 * k, S + 6 * k, S + 7 * k, S + 8 * k, S + 9 * k, ns, d_x, d_y, d_z,
                              d_mass, d_ax, d_ay, d_az, n, parameters.gravity, stream, t ? &ms2 : nullptr),
              "starForces");
    time_computeForces[step] = ms + ms2;
    if (t) Log(LogLevel::TIME) << "\tCompute forces: " << ms + ms2 << " ms";
    check(nbx_update(ctx, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_ax, d_ay, d_az, n, parameters.timestep,
                     parameters.dampening, stream, t ? &ms : nullptr),
          "update");
    ms2 = 0.f;
    if (ns > 0)
        check(nbx_update(ctx, S, S + k, S + 2 * k, S + 3 * k, S + 4 * k, S + 5 * k, S + 7 * k, S + 8 * k, S + 9 * k, ns,
                         parameters.timestep, parameters.dampening, stream, t ? &ms2 : nullptr),
              "update suns");
    time_update[step] = ms + ms2;
    if (t) Log(LogLevel::TIME) << "\tUpdate particles: " << ms + ms2 << " ms";
    float copy_ms = 0.f;
    if (parameters.copyToHost) {
        hipEvent_t c0, c1;
        hipEventCreate(&c0);
        hipEventCreate(&c1);
        hipEventRecord(c0, stream);
        hipMemcpyAsync(h_x.data(), d_x, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipMemcpyAsync(h_y.data(), d_y, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipMemcpyAsync(h_z.data(), d_z, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipMemcpyAsync(h_vx.data(), d_vx, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipMemcpyAsync(h_vy.data(), d_vy, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipMemcpyAsync(h_vz.data(), d_vz, 4 * (size_t)n, hipMemcpyDeviceToHost, stream);
        hipEventRecord(c1, stream);
        hipEventSynchronize(c1);
        hipEventElapsedTime(&copy_ms, c0, c1);
        hipEventDestroy(c0);
        hipEventDestroy(c1);
    }
    time_copyDeviceToHost[step] = copy_ms;
    if (t) Log(LogLevel::TIME) << "\tCopying to host: " << copy_ms << " ms";
    hipEventRecord(b, stream);
    hipEventSynchronize(b);
    hipEventElapsedTime(&ms, a, b);
    time_all[step] = ms;
    hipEventDestroy(a);
    hipEventDestroy(b);
    check(nbx_check(ctx, stream), "step");
    Log(LogLevel::TIME) << "Elapsed time for step " << step << " : " << ms << " ms";
}

float BarnesHut::getSystemSize() const  // src/BarnesHut.cu:418-449
{
    float s = 0.f;
    for (int i = 0; i < numParticles; ++i)
        s = std::fmax(s, std::fmax(std::fabs(h_x[i]), std::fmax(std::fabs(h_y[i]), std::fabs(h_z[i]))));
    return s;
}

}  // namespace nbx
