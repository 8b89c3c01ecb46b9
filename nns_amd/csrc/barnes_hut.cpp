// barnes_hut.cpp -- the reference's BarnesHut simulation object
// (3D/CUDA/CUDA_NBody/src/BarnesHut.cu:12-301) over the nbx C ABI.
#include "barnes_hut.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace nbx {

namespace {
template <class T> T *dmalloc(size_t count)
{
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    return static_cast<T *>(p);
}
}  // namespace

BarnesHut::BarnesHut(const SimulationParameters &p) : parameters(p)
{
    numParticles = p.numberOfParticles;
    numNodes = 2 * numParticles + 12000;  // src/BarnesHut.cu:18
    const size_t n = numParticles, m = numNodes;
    for (auto *v : {&h_x, &h_y, &h_z, &h_vx, &h_vy, &h_vz, &h_mass}) v->assign(n, 0.f);
    for (auto *v : {&time_resetArrays, &time_computeBoundingBox, &time_buildTree, &time_centreOfMass, &time_sort,
                    &time_computeForces, &time_update, &time_copyDeviceToHost, &time_all})
        v->assign(p.iterations, 0.f);
    int rc;
    if (p.model == "plummer")
        rc = nbx_ic_plummer(numParticles, p.seed, h_x.data(), h_y.data(), h_z.data(), h_vx.data(), h_vy.data(),
                            h_vz.data(), h_mass.data());
    else if (p.model == "star_disk")
        rc = nbx_ic_star_disk(numParticles, p.seed, h_x.data(), h_y.data(), h_z.data(), h_vx.data(), h_vy.data(),
                              h_vz.data(), h_mass.data());
    else
        rc = nbx_ic_binary_disk(numParticles, p.seed, h_x.data(), h_y.data(), h_z.data(), h_vx.data(),
                                h_vy.data(), h_vz.data(), h_mass.data());
    if (rc != NBX_OK) throw std::runtime_error("initial conditions: bad parameters");
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
    hipMemcpy(d_x, h_x.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_y, h_y.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_z, h_z.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_vx, h_vx.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_vy, h_vy.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_vz, h_vz.data(), 4 * n, hipMemcpyHostToDevice);
    hipMemcpy(d_mass, h_mass.data(), 4 * n, hipMemcpyHostToDevice);
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
                    (void *)d_max_x, (void *)d_min_y, (void *)d_max_y, (void *)d_min_z, (void *)d_max_z})
        if (d) hipFree(d);
    if (stream) hipStreamDestroy(stream);
}

void BarnesHut::check(int rc, const char *what)
{
    if (rc != NBX_OK) throw std::runtime_error(std::string(what) + ": " + nbx_last_error(ctx));
}

void BarnesHut::update(int step)
{
    const int n = numParticles, m = numNodes;
    const bool t = parameters.timeKernels;
    float ms = 0.f;
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, stream);
    check(nbx_reset_arrays(ctx, d_mutex, d_x, d_y, d_z, d_mass, d_count, d_start, d_sorted, d_child, d_index, d_min_x,
                           d_max_x, d_min_y, d_max_y, d_min_z, d_max_z, n, m, stream, t ? &ms : nullptr),
          "resetArrays");
    time_resetArrays[step] = ms;
    check(nbx_compute_bounding_box(ctx, d_mutex, d_x, d_y, d_z, d_min_x, d_max_x, d_min_y, d_max_y, d_min_z,
                                   d_max_z, n, stream, t ? &ms : nullptr),
          "computeBoundingBox");
    time_computeBoundingBox[step] = ms;
    check(nbx_build_tree(ctx, d_x, d_y, d_z, d_mass, d_count, d_start, d_child, d_index, d_min_x, d_max_x, d_min_y,
                         d_max_y, d_min_z, d_max_z, n, m, stream, t ? &ms : nullptr),
          "buildTree");
    time_buildTree[step] = ms;
    check(nbx_centre_of_mass(ctx, d_x, d_y, d_z, d_mass, d_index, n, stream, t ? &ms : nullptr), "centreOfMass");
    time_centreOfMass[step] = ms;
    check(nbx_sort(ctx, d_count, d_start, d_sorted, d_child, d_index, n, stream, t ? &ms : nullptr), "sort");
    time_sort[step] = ms;
    check(nbx_compute_forces(ctx, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_ax, d_ay, d_az, d_mass, d_sorted, d_child,
                             d_min_x, d_max_x, n, parameters.gravity, stream, t ? &ms : nullptr),
          "computeForces");
    time_computeForces[step] = ms;
    check(nbx_update(ctx, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_ax, d_ay, d_az, n, parameters.timestep,
                     parameters.dampening, stream, t ? &ms : nullptr),
          "update");
    time_update[step] = ms;
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
    hipEventRecord(b, stream);
    hipEventSynchronize(b);
    hipEventElapsedTime(&ms, a, b);
    time_all[step] = ms;
    hipEventDestroy(a);
    hipEventDestroy(b);
    check(nbx_check(ctx, stream), "step");
}

float BarnesHut::getSystemSize() const  // src/BarnesHut.cu:418-449
{
    float s = 0.f;
    for (int i = 0; i < numParticles; ++i)
        s = std::fmax(s, std::fmax(std::fabs(h_x[i]), std::fmax(std::fabs(h_y[i]), std::fabs(h_z[i]))));
    return s;
}

}  // namespace nbx
