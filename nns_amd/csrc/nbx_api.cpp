// nbx_api.cpp -- the C ABI declared in include/nbx.h.
//
// Host side of the drop-in for the reference's KernelsWrapper
// (3D/CUDA/CUDA_NBody/src/KernelsWrapper.cu:80-208): argument checks, stage
// ordering, launches on the caller's stream, optional hipEvent timing, and the
// device-side error latch.  No stage allocates memory; nbx_create does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <new>
#include <string>

#include "nbx_internal.h"

using nbx::DeviceState;

namespace nbx {

hipError_t copy_cell_sums(const double4 *sum64, bool com, double *dst, size_t count)
{
    const size_t ncell = com ? (count + 2) / 3 : count;
    std::vector<double4> v(ncell ? ncell : 1);
    const hipError_t e = hipMemcpy(v.data(), sum64, sizeof(double4) * ncell, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (size_t k = 0; k < count; ++k) {
        const double4 &s = v[com ? k / 3 : k];
        if (!com) {
            dst[k] = s.x;
        } else {
            const int a = (int)(k % 3);
            const double mk = a == 0 ? s.y : a == 1 ? s.z : s.w;
            dst[k] = s.x != 0.0 ? mk / s.x : 0.0;
        }
    }
    return hipSuccess;
}

}  // namespace nbx

namespace {

thread_local std::string g_create_error;

int fail(nbx_context *c, int code, const std::string &msg)
{
    if (c) c->last_error = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(nbx_context *c, hipError_t e, const char *what)
{
    return fail(c, NBX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NBX_HIP(ctx, call)                                  \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

bool params_ok(const nbx_params *p)
{
    return p && p->theta >= 0.f && p->softening >= 0.f && p->timestep == p->timestep &&
           p->gravity == p->gravity && p->dampening == p->dampening &&
           (p->block_size == 0 || (p->block_size >= 64 && p->block_size <= 1024 && p->block_size % 64 == 0)) &&
           (p->force_waves == 0 || p->force_waves == 2 || p->force_waves == 4 || p->force_waves == 8);
}

struct StageTimer {
    nbx_context *c;
    hipStream_t s;
    float *out;
    StageTimer(nbx_context *c_, hipStream_t s_, float *out_) : c(c_), s(s_), out(out_)
    {
        if (out) hipEventRecord(c->ev0, s);
    }
    int finish()
    {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
        if (out) {
            hipEventRecord(c->ev1, s);
            e = hipEventSynchronize(c->ev1);
            if (e != hipSuccess) return hip_fail(c, e, "hipEventSynchronize");
            hipEventElapsedTime(out, c->ev0, c->ev1);
        }
        return NBX_OK;
    }
};

template <class T> hipError_t dalloc(T **p, size_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (count ? count : 1));
}

int check_sizes(nbx_context *c, int n, int m)
{
    if (n < 0 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [0, n_max]");
    if (m < n + 1 || m > c->m_max) return fail(c, NBX_ERR_INVALID, "m must satisfy n < m <= m_max");
    return NBX_OK;
}

}  // namespace

void nbx::set_create_error(const std::string &msg) { g_create_error = msg; }

extern "C" {

void nbx_default_params(nbx_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->theta = 0.5f;
    p->softening = 1e-3f;
    p->gravity = 1.f;
    p->timestep = 1e-3f;
    p->dampening = 1.f;
    p->kahan = 1;
    p->block_size = 256;
    p->force_waves = 4;
}

int nbx_create(const nbx_params *p, int n_max, int m_max, int device, nbx_context **out)
{
    if (!out) return fail(nullptr, NBX_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!params_ok(p)) return fail(nullptr, NBX_ERR_INVALID, "invalid parameters");
    if (n_max < 1 || m_max < n_max + 1) return fail(nullptr, NBX_ERR_INVALID, "need 1 <= n_max < m_max");
    if (n_max > nbx::kMaxBodies) return fail(nullptr, NBX_ERR_INVALID, "n_max above the maximum of 48,805,888 bodies");
    if (m_max > nbx::kMaxNodes) return fail(nullptr, NBX_ERR_INVALID, "m_max above the maximum of 134,217,728 nodes");
    nbx_context *c = new (std::nothrow) nbx_context();
    if (!c) return fail(nullptr, NBX_ERR_INVALID, "out of host memory");
    c->p = *p;
    c->force_waves = p->force_waves ? p->force_waves : 4;
    c->device = device;
    c->n_max = n_max;
    c->m_max = m_max;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, NBX_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    const size_t n = (size_t)n_max, m = (size_t)m_max;
    c->bbox_blocks = 1024;
    c->hist_tiles = (n_max + nbx::kSortTile - 1) / nbx::kSortTile;
    bool ok = dalloc(&c->st, 1) == hipSuccess && dalloc(&c->bbox_part, 6 * (size_t)c->bbox_blocks) == hipSuccess &&
              dalloc(&c->keys, n) == hipSuccess && dalloc(&c->skeys, n) == hipSuccess &&
              dalloc(&c->kbuf, n) == hipSuccess && dalloc(&c->perm, n) == hipSuccess &&
              dalloc(&c->vbuf, n) == hipSuccess && dalloc(&c->pos, n) == hipSuccess &&
              dalloc(&c->hist, (size_t)c->hist_tiles * nbx::kRadix + nbx::kRadix) == hipSuccess &&
              dalloc(&c->cell_level, m) == hipSuccess && dalloc(&c->sum64, m) == hipSuccess &&
              dalloc(&c->nodes4, m) == hipSuccess &&
              dalloc(&c->fnode, m + 8) == hipSuccess && dalloc(&c->fmeta, m + 8) == hipSuccess &&
              dalloc(&c->boff, m + 1) == hipSuccess && dalloc(&c->bsum, m / nbx::kOffRun + 2) == hipSuccess &&
              dalloc(&c->froot4, 1) == hipSuccess && dalloc(&c->frootm, 1) == hipSuccess &&
              dalloc(&c->pblk, m) == hipSuccess && dalloc(&c->bpar, n) == hipSuccess && dalloc(&c->counts, n) == hipSuccess &&
              dalloc(&c->acc_scratch, 3 * n) == hipSuccess && dalloc(&c->bodies4, n) == hipSuccess &&
              dalloc(&c->star_part, (size_t)nbx::kStarPartBlocks * NBX_MAX_STARS * 3) == hipSuccess &&
              dalloc(&c->walk, 2 * ((n + 63) / 64)) == hipSuccess && dalloc(&c->gcost, (n + 63) / 64) == hipSuccess &&
              dalloc(&c->gkey, n) == hipSuccess && dalloc(&c->gkey2, n) == hipSuccess &&
              dalloc(&c->gorder, n) == hipSuccess &&
              hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
    for (int k = 0; ok && k < 8; ++k) ok = hipEventCreate(&c->stage_ev[k]) == hipSuccess;
    if (ok) ok = nbx::alloc_tree_scratch(c->ts, n_max, 1, (n_max + nbx::kSubtreeMax - 1) / nbx::kSubtreeMax);
    if (ok) ok = hipMemset(c->st, 0, sizeof(DeviceState)) == hipSuccess &&
                 hipMemset(c->gcost, 0, sizeof(int32_t) * ((n + 63) / 64)) == hipSuccess;
    if (ok) ok = hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        nbx_destroy(c);
        return fail(nullptr, NBX_ERR_HIP, "device allocation failed");
    }
    *out = c;
    return NBX_OK;
}

int nbx_destroy(nbx_context *c)
{
    if (!c) return NBX_OK;
    hipSetDevice(c->device);
    hipDeviceSynchronize();
    void *bufs[] = {c->st, c->bbox_part, c->keys, c->skeys, c->kbuf, c->perm, c->vbuf, c->pos, c->hist,
                    c->cell_level, c->sum64, c->nodes4, c->fnode, c->fmeta, c->boff, c->bsum, c->froot4, c->frootm, c->pblk, c->bpar, c->counts,
                    c->acc_scratch, c->star_part, c->walk, c->gcost, c->bodies4, c->gkey, c->gkey2,
                    c->gorder};
    for (void *b : bufs)
        if (b) hipFree(b);
    nbx::free_tree_scratch(c->ts);
    nbx::free_vjp_scratch(c);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (auto &ev : c->stage_ev)
        if (ev) hipEventDestroy(ev);
    delete c;
    return NBX_OK;
}

int nbx_set_params(nbx_context *c, const nbx_params *p)
{
    if (!c) return NBX_ERR_INVALID;
    if (!params_ok(p)) return fail(c, NBX_ERR_INVALID, "invalid parameters");
    c->p = *p;
    if (p->force_waves) c->force_waves = p->force_waves;  // 0 keeps the current (possibly autotuned) choice
    return NBX_OK;
}

int nbx_get_params(const nbx_context *c, nbx_params *p)
{
    if (!c || !p) return NBX_ERR_INVALID;
    *p = c->p;
    p->force_waves = c->force_waves;
    return NBX_OK;
}

const char *nbx_last_error(const nbx_context *c)
{
    return c ? c->last_error.c_str() : g_create_error.c_str();
}

int nbx_check(nbx_context *c, void *stream)
{
    if (!c) return NBX_ERR_INVALID;
    hipSetDevice(c->device);
    NBX_HIP(c, hipStreamSynchronize((hipStream_t)stream));
    // e[0] err: this step's latch; e[1] sticky: unreported errors of earlier
    // steps (moved there by the reset stage); e[2] seen.  Reported here once:
    // sticky is cleared and this step's bits are marked seen, so that the next
    // reset does not report them again.
    int e[3] = {0, 0, 0};
    static_assert(offsetof(DeviceState, sticky) == offsetof(DeviceState, err) + sizeof(int) &&
                      offsetof(DeviceState, seen) == offsetof(DeviceState, err) + 2 * sizeof(int),
                  "err, sticky and seen are adjacent");
    NBX_HIP(c, hipMemcpy(e, &c->st->err, 3 * sizeof(int), hipMemcpyDeviceToHost));
    if (e[1] || e[2] != e[0]) {
        const int upd[2] = {0, e[0]};
        NBX_HIP(c, hipMemcpy(&c->st->sticky, upd, 2 * sizeof(int), hipMemcpyHostToDevice));
    }
    const int err = e[0] | e[1];
    if (err & nbx::kErrNonFinite) return fail(c, NBX_ERR_NONFINITE, "non-finite position or zero cell mass");
    if (err & nbx::kErrCapacity) return fail(c, NBX_ERR_CAPACITY, "tree needs more than m - n cells");
    if (err & nbx::kErrBucket) return fail(c, NBX_ERR_BUCKET, "more than 8 bodies share one key");
    if (err & nbx::kErrStack) return fail(c, NBX_ERR_STACK, "traversal stack overflow");
    if (err & nbx::kErrStall) return fail(c, NBX_ERR_HIP, "top summary: its workgroups did not run together");
    return NBX_OK;
}

int nbx_reset_arrays(nbx_context *c, int *mutex, float *x, float *y, float *z, float *mass, int *count,
                     int *start, int *sorted, int *child, int *index, float *minX, float *maxX, float *minY,
                     float *maxY, float *minZ, float *maxZ, int n, int m, void *stream, float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (int r = check_sizes(c, n, m)) return r;
    if (!mutex || !x || !y || !z || !mass || !count || !start || !sorted || !child || !index || !minX || !maxX ||
        !minY || !maxY || !minZ || !maxZ)
        return fail(c, NBX_ERR_INVALID, "null array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_reset(c, s, mutex, x, y, z, mass, count, start, sorted, child, index, minX, maxX, minY, maxY,
                      minZ, maxZ, n, m);
    c->stage = 0;
    c->fresh = true;
    return t.finish();
}

int nbx_compute_bounding_box(nbx_context *c, int *mutex, float *x, float *y, float *z, float *minX, float *maxX,
                             float *minY, float *maxY, float *minZ, float *maxZ, int n, void *stream,
                             float *elapsed_ms)
{
    (void)mutex;  // the reference's global merge lock; the merge here is lock-free
    if (!c) return NBX_ERR_INVALID;
    if (n < 1 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [1, n_max]");
    if (!x || !y || !z || !minX || !maxX || !minY || !maxY || !minZ || !maxZ)
        return fail(c, NBX_ERR_INVALID, "null array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_bbox(c, s, x, y, z, minX, maxX, minY, maxY, minZ, maxZ, n);
    c->stage = 1;
    c->bbox_x = x; c->bbox_y = y; c->bbox_z = z; c->bbox_n = n;
    return t.finish();
}

int nbx_build_tree(nbx_context *c, float *x, float *y, float *z, float *mass, int *count, int *start,
                   int *child, int *index, float *minX, float *maxX, float *minY, float *maxY, float *minZ,
                   float *maxZ, int n, int m, void *stream, float *elapsed_ms)
{
    (void)minX; (void)maxX; (void)minY; (void)maxY; (void)minZ; (void)maxZ;  // extent is kept in FP64
    if (!c) return NBX_ERR_INVALID;
    if (int r = check_sizes(c, n, m)) return r;
    if (n < 1) return fail(c, NBX_ERR_INVALID, "n must be >= 1");
    if (!x || !y || !z || !mass || !count || !start || !child || !index)
        return fail(c, NBX_ERR_INVALID, "null array");
    if (c->stage < 1) return fail(c, NBX_ERR_ORDER, "nbx_build_tree needs nbx_compute_bounding_box first");
    if (!c->fresh) return fail(c, NBX_ERR_ORDER, "nbx_build_tree needs nbx_reset_arrays first");
    if (n != c->bbox_n || x != c->bbox_x || y != c->bbox_y || z != c->bbox_z)
        return fail(c, NBX_ERR_INVALID, "bodies differ from those of the last nbx_compute_bounding_box");
    c->fresh = false;
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    c->x = x; c->y = y; c->z = z; c->mass = mass;
    c->child = child; c->count = count; c->start = start; c->index = index;
    c->n = n; c->m = m;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_keys(c, s, x, y, z, mass, n, NBX_KEY_MORTON, c->keys);
    nbx::launch_radix_sort(c, s, c->keys, n, c->skeys, c->perm, c->kbuf, c->vbuf, c->pos, nbx::kMortonSortBits);
    nbx::launch_build(c, s, n, m, child, index);
    c->stage = 2;
    return t.finish();
}

int nbx_centre_of_mass(nbx_context *c, float *x, float *y, float *z, float *mass, int *index, int n,
                       void *stream, float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 2) return fail(c, NBX_ERR_ORDER, "nbx_centre_of_mass needs nbx_build_tree first");
    if (x != c->x || y != c->y || z != c->z || mass != c->mass || index != c->index || n != c->n)
        return fail(c, NBX_ERR_INVALID, "arrays differ from those given to nbx_build_tree");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_summarize(c, s);
    c->stage = 3;
    return t.finish();
}

int nbx_sort(nbx_context *c, int *count, int *start, int *sorted, int *child, int *index, int n, void *stream,
             float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 3) return fail(c, NBX_ERR_ORDER, "nbx_sort needs nbx_centre_of_mass first");
    if (count != c->count || start != c->start || child != c->child || index != c->index || n != c->n)
        return fail(c, NBX_ERR_INVALID, "arrays differ from those given to nbx_build_tree");
    if (!sorted) return fail(c, NBX_ERR_INVALID, "null array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_layout(c, s, sorted);
    c->stage = 4;
    return t.finish();
}

int nbx_compute_forces(nbx_context *c, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax,
                       float *ay, float *az, float *mass, int *sorted, int *child, float *minX, float *maxX,
                       int n, float g, void *stream, float *elapsed_ms)
{
    (void)vx; (void)vy; (void)vz; (void)sorted; (void)minX; (void)maxX;
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 4) return fail(c, NBX_ERR_ORDER, "nbx_compute_forces needs nbx_sort first");
    if (x != c->x || y != c->y || z != c->z || mass != c->mass || child != c->child || n != c->n)
        return fail(c, NBX_ERR_INVALID, "arrays differ from those given to nbx_build_tree");
    if (!ax || !ay || !az) return fail(c, NBX_ERR_INVALID, "null array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_force(c, s, ax, ay, az, g, nullptr);
    c->stage = 5;
    return t.finish();
}

int nbx_update(nbx_context *c, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax, float *ay,
               float *az, int n, float dt, float d, void *stream, float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (n < 0 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [0, n_max]");
    if (!x || !y || !z || !vx || !vy || !vz || !ax || !ay || !az) return fail(c, NBX_ERR_INVALID, "null array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_update(s, c->p.block_size, c->st, x, y, z, vx, vy, vz, ax, ay, az, n, dt, d);
    return t.finish();
}

int nbx_star_forces(nbx_context *c, const float *sx, const float *sy, const float *sz, const float *smass,
                    float *sax, float *say, float *saz, int ns, const float *x, const float *y, const float *z,
                    const float *mass, float *ax, float *ay, float *az, int n, float g, void *stream,
                    float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (ns < 0 || ns > NBX_MAX_STARS) return fail(c, NBX_ERR_INVALID, "ns out of range [0, NBX_MAX_STARS]");
    if (n < 0 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [0, n_max]");
    if (ns > 0 && (!sx || !sy || !sz || !smass || !sax || !say || !saz))
        return fail(c, NBX_ERR_INVALID, "null star array");
    if (n > 0 && (!x || !y || !z || !mass || !ax || !ay || !az)) return fail(c, NBX_ERR_INVALID, "null array");
    if (!std::isfinite(g)) return fail(c, NBX_ERR_INVALID, "g must be finite");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    if (ns > 0) nbx::launch_star_forces(c, s, sx, sy, sz, smass, sax, say, saz, ns, x, y, z, mass, ax, ay, az, n, g);
    return t.finish();
}

int nbx_compute_keys(nbx_context *c, const float *x, const float *y, const float *z, int n, int key_type,
                     uint64_t *keys, void *stream, float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (n < 0 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [0, n_max]");
    if (!x || !y || !z || !keys) return fail(c, NBX_ERR_INVALID, "null array");
    if (key_type != NBX_KEY_MORTON && key_type != NBX_KEY_HILBERT)
        return fail(c, NBX_ERR_INVALID, "unknown key type");
    if (c->stage < 1) return fail(c, NBX_ERR_ORDER, "nbx_compute_keys needs nbx_compute_bounding_box first");
    if (n != c->bbox_n || x != c->bbox_x || y != c->bbox_y || z != c->bbox_z)
        return fail(c, NBX_ERR_INVALID, "bodies differ from those of the last nbx_compute_bounding_box");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_keys(c, s, x, y, z, nullptr, n, key_type, keys);
    return t.finish();
}

int nbx_sort_keys(nbx_context *c, const uint64_t *keys, int n, uint64_t *sorted_keys, int32_t *perm, void *stream,
                  float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (n < 0 || n > c->n_max) return fail(c, NBX_ERR_INVALID, "n out of range [0, n_max]");
    if (!keys || !sorted_keys || !perm) return fail(c, NBX_ERR_INVALID, "null array");
    if ((const void *)keys == (void *)sorted_keys) return fail(c, NBX_ERR_INVALID, "keys and sorted_keys alias");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_radix_sort(c, s, keys, n, sorted_keys, perm, c->kbuf, c->vbuf, nullptr, 64);
    return t.finish();
}

int nbx_count_interactions(nbx_context *c, int n, int64_t *counts, void *stream)
{
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 4) return fail(c, NBX_ERR_ORDER, "nbx_count_interactions needs nbx_sort first");
    if (n != c->n || !counts) return fail(c, NBX_ERR_INVALID, "n or counts invalid");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    // the counting variant also writes accelerations; they go to context scratch
    float *scratch = c->acc_scratch;
    StageTimer t(c, s, nullptr);
    nbx::launch_force(c, s, scratch, scratch + n, scratch + 2 * (size_t)n, 1.f, counts);
    return t.finish();
}

int nbx_walk_stats(nbx_context *c, int64_t *out, void *stream)
{
    if (!c || !out) return NBX_ERR_INVALID;
    if (c->stage < 4) return fail(c, NBX_ERR_ORDER, "nbx_walk_stats needs nbx_sort first");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    const int n = c->n, ng = (n + 63) / 64;
    float *scratch = c->acc_scratch;
    nbx::launch_force(c, s, scratch, scratch + n, scratch + 2 * (size_t)n, 1.f, c->counts, c->walk);
    NBX_HIP(c, hipGetLastError());
    NBX_HIP(c, hipStreamSynchronize(s));
    std::vector<int64_t> cnt(n), wk(2 * (size_t)ng);
    NBX_HIP(c, hipMemcpy(cnt.data(), c->counts, 8 * (size_t)n, hipMemcpyDeviceToHost));
    NBX_HIP(c, hipMemcpy(wk.data(), c->walk, 16 * (size_t)ng, hipMemcpyDeviceToHost));
    int64_t inter = 0, steps = 0, frames = 0;
    for (int64_t v : cnt) inter += v;
    for (int k = 0; k < ng; ++k) {
        steps += wk[2 * (size_t)k];
        frames += wk[2 * (size_t)k + 1];
    }
    out[0] = inter;
    out[1] = steps;
    out[2] = frames;
    out[3] = ng;
    return NBX_OK;
}

int nbx_force_vjp(nbx_context *c, const float *ux, const float *uy, const float *uz, float g, float *gx, float *gy,
                  float *gz, float *gm, void *stream, float *elapsed_ms)
{
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 4) return fail(c, NBX_ERR_ORDER, "nbx_force_vjp needs nbx_sort first");
    if (!ux || !uy || !uz || !gx || !gy || !gz || !gm) return fail(c, NBX_ERR_INVALID, "null array");
    if (!std::isfinite(g)) return fail(c, NBX_ERR_INVALID, "g must be finite");
    hipSetDevice(c->device);
    if (!c->adj_tgt && !nbx::alloc_vjp_scratch(c)) return fail(c, NBX_ERR_HIP, "adjoint scratch allocation failed");
    hipStream_t s = (hipStream_t)stream;
    StageTimer t(c, s, elapsed_ms);
    nbx::launch_vjp(c, s, ux, uy, uz, gx, gy, gz, gm, g);
    return t.finish();
}

int nbx_step(nbx_context *c, const nbx_arrays *a, void *stream, float *stage_ms)
{
    if (!c || !a) return NBX_ERR_INVALID;
    if (int r = check_sizes(c, a->n, a->m)) return r;
    if (a->n < 1) return fail(c, NBX_ERR_INVALID, "n must be >= 1");
    const int ns = a->nstars;
    if (ns < 0 || ns > NBX_MAX_STARS) return fail(c, NBX_ERR_INVALID, "nstars out of range [0, NBX_MAX_STARS]");
    if (ns > 0 && (!a->sx || !a->sy || !a->sz || !a->smass || !a->svx || !a->svy || !a->svz || !a->sax || !a->say ||
                   !a->saz))
        return fail(c, NBX_ERR_INVALID, "null star array");
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    const bool tm = stage_ms != nullptr;
    auto mark = [&](int k) {
        if (tm) hipEventRecord(c->stage_ev[k], s);
    };
    mark(0);
    int r = nbx_reset_arrays(c, a->mutex, a->x, a->y, a->z, a->mass, a->count, a->start, a->sorted, a->child,
                             a->index, a->minX, a->maxX, a->minY, a->maxY, a->minZ, a->maxZ, a->n, a->m, stream,
                             nullptr);
    mark(1);
    if (!r) r = nbx_compute_bounding_box(c, a->mutex, a->x, a->y, a->z, a->minX, a->maxX, a->minY, a->maxY,
                                         a->minZ, a->maxZ, a->n, stream, nullptr);
    mark(2);
    if (!r) r = nbx_build_tree(c, a->x, a->y, a->z, a->mass, a->count, a->start, a->child, a->index, a->minX,
                               a->maxX, a->minY, a->maxY, a->minZ, a->maxZ, a->n, a->m, stream, nullptr);
    mark(3);
    if (!r) r = nbx_centre_of_mass(c, a->x, a->y, a->z, a->mass, a->index, a->n, stream, nullptr);
    mark(4);
    if (!r) r = nbx_sort(c, a->count, a->start, a->sorted, a->child, a->index, a->n, stream, nullptr);
    mark(5);
    if (!r) r = nbx_compute_forces(c, a->x, a->y, a->z, a->vx, a->vy, a->vz, a->ax, a->ay, a->az, a->mass,
                                   a->sorted, a->child, a->minX, a->maxX, a->n, c->p.gravity, stream, nullptr);
    if (!r && ns > 0)
        r = nbx_star_forces(c, a->sx, a->sy, a->sz, a->smass, a->sax, a->say, a->saz, ns, a->x, a->y, a->z, a->mass,
                            a->ax, a->ay, a->az, a->n, c->p.gravity, stream, nullptr);
    mark(6);
    if (!r) r = nbx_update(c, a->x, a->y, a->z, a->vx, a->vy, a->vz, a->ax, a->ay, a->az, a->n, c->p.timestep,
                           c->p.dampening, stream, nullptr);
    if (!r && ns > 0) {  // the stars move with the bodies (Interaction.cpp:219-234)
        nbx::launch_update(s, 64, c->st, a->sx, a->sy, a->sz, a->svx, a->svy, a->svz, a->sax, a->say, a->saz, ns,
                           c->p.timestep, c->p.dampening);
        if (hipGetLastError() != hipSuccess) r = fail(c, NBX_ERR_HIP, "kernel launch: star update");
    }
    mark(7);
    if (r) return r;
    if (tm) {
        NBX_HIP(c, hipEventSynchronize(c->stage_ev[7]));
        for (int k = 0; k < 7; ++k) hipEventElapsedTime(&stage_ms[k], c->stage_ev[k], c->stage_ev[k + 1]);
    }
    return NBX_OK;
}

int nbx_copy_buffer(nbx_context *c, int which, void *dst, int64_t count, void *stream)
{
    if (!c || !dst || count < 0) return NBX_ERR_INVALID;
    hipSetDevice(c->device);
    const void *src = nullptr;
    size_t elem = 0, cap = 0;
    switch (which) {
    case NBX_BUF_KEYS: src = c->keys; elem = 8; cap = c->n_max; break;
    case NBX_BUF_SORTED_KEYS: src = c->skeys; elem = 8; cap = c->n_max; break;
    case NBX_BUF_PERM: src = c->perm; elem = 4; cap = c->n_max; break;
    case NBX_BUF_CELL_LEVEL: src = c->cell_level; elem = 4; cap = c->m_max; break;
    case NBX_BUF_CELL_MASS64: elem = 8; cap = c->m_max; break;
    case NBX_BUF_CELL_COM64: elem = 8; cap = 3 * (size_t)c->m_max; break;
    case NBX_BUF_ROOT: src = c->st->root64; elem = 8; cap = 4; break;
    case NBX_BUF_EXTENT: src = c->st->ext; elem = 8; cap = 4; break;
    default: return fail(c, NBX_ERR_INVALID, "unknown buffer");
    }
    if ((size_t)count > cap) return fail(c, NBX_ERR_INVALID, "count exceeds buffer");
    NBX_HIP(c, hipStreamSynchronize((hipStream_t)stream));
    if (which == NBX_BUF_CELL_MASS64 || which == NBX_BUF_CELL_COM64)
        NBX_HIP(c, nbx::copy_cell_sums(c->sum64, which == NBX_BUF_CELL_COM64, static_cast<double *>(dst),
                                       (size_t)count));
    else
        NBX_HIP(c, hipMemcpy(dst, src, elem * (size_t)count, hipMemcpyDeviceToHost));
    return NBX_OK;
}

}  // extern "C"
