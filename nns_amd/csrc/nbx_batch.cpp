// nbx_batch.cpp -- batched ensembles and the force-kernel autotuner (C ABI of
// include/nbx.h).  No reference counterpart: the reference runs one system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "nbx_internal.h"

namespace {

thread_local std::string g_batch_error;

int bfail(nbx_batch *b, int code, const std::string &msg)
{
    if (b) b->last_error = msg;
    else g_batch_error = msg;
    return code;
}

template <class T> bool dalloc(T **p, size_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (count ? count : 1)) == hipSuccess;
}

int finish(nbx_batch *b)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, NBX_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return NBX_OK;
}

// Stage sequence of one step; marks are recorded only when events are given.
void enqueue_step(nbx_batch *b, hipStream_t s, hipEvent_t *ev)
{
    auto mark = [&](int k) {
        if (ev) hipEventRecord(ev[k], s);
    };
    mark(0);
    nbx::launch_batch_reset(b, s);
    mark(1);
    nbx::launch_batch_bbox(b, s);
    mark(2);
    nbx::launch_batch_keys(b, s);
    nbx::launch_batch_sort(b, s);
    nbx::launch_batch_build(b, s);
    mark(3);
    nbx::launch_batch_summarize(b, s);
    mark(4);
    nbx::launch_batch_layout(b, s);
    mark(5);
    nbx::launch_batch_force(b, s, nullptr, b->force_waves);
    mark(6);
    nbx::launch_update(s, b->p.block_size, b->st, b->x, b->y, b->z, b->vx, b->vy, b->vz, b->ax, b->ay, b->az, b->N,
                       b->p.timestep, b->p.dampening);
    mark(7);
}

void drop_graph(nbx_batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

bool bound(const nbx_batch *b) { return b->x && b->y && b->z && b->vx && b->vy && b->vz && b->ax && b->mass; }

}  // namespace

extern "C" {

int nbx_batch_create(const nbx_params *p, int nsys, const int *sys_n, const float *sys_theta, int device,
                     nbx_batch **out)
{
    if (!out) return bfail(nullptr, NBX_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!p || nsys < 1 || !sys_n) return bfail(nullptr, NBX_ERR_INVALID, "need params, nsys >= 1 and sys_n");
    if (nsys > 65535) return bfail(nullptr, NBX_ERR_INVALID, "at most 65535 systems");
    if (p->force_waves != 0 && p->force_waves != 2 && p->force_waves != 4 && p->force_waves != 8)
        return bfail(nullptr, NBX_ERR_INVALID, "force_waves must be 0, 2, 4 or 8");
    int64_t total = 0, samples = 0;
    for (int s = 0; s < nsys; ++s) {
        if (sys_n[s] < 1) return bfail(nullptr, NBX_ERR_INVALID, "every system needs >= 1 body");
        total += sys_n[s];
        samples += (sys_n[s] + nbx::kSubtreeMax - 1) / nbx::kSubtreeMax;
    }
    if (samples > nbx::kMaxSamples)
        return bfail(nullptr, NBX_ERR_INVALID, "ensemble above the maximum of 47,662 samples (sum of ceil(n_s / 1024))");
    const int64_t m64 = 2 * total + nsys + 12000;
    if (m64 > nbx::kMaxNodes) return bfail(nullptr, NBX_ERR_INVALID, "ensemble above the maximum of 134,217,728 nodes");
    if (sys_theta)
        for (int s = 0; s < nsys; ++s)
            if (!(sys_theta[s] >= 0.f)) return bfail(nullptr, NBX_ERR_INVALID, "invalid theta");
    if (hipSetDevice(device) != hipSuccess) return bfail(nullptr, NBX_ERR_HIP, "hipSetDevice failed");
    nbx_batch *b = new (std::nothrow) nbx_batch();
    if (!b) return bfail(nullptr, NBX_ERR_INVALID, "out of host memory");
    b->p = *p;
    b->device = device;
    b->S = nsys;
    b->N = (int)total;
    b->m = (int)m64;
    b->force_waves = p->force_waves == 2 || p->force_waves == 8 ? p->force_waves : 4;
    std::vector<int32_t> off(nsys + 1, 0), sys_of(total), samp(nsys + 1, 0);
    std::vector<float> th(nsys, p->theta);
    std::vector<int4> groups, chunks;
    std::vector<int32_t> choff(nsys + 1, 0);
    for (int s = 0; s < nsys; ++s) {
        off[s + 1] = off[s] + sys_n[s];
        for (int f = off[s]; f < off[s + 1]; f += nbx::kBboxChunk)
            chunks.push_back(make_int4(s, f, std::min(f + nbx::kBboxChunk, off[s + 1]), 0));
        choff[s + 1] = (int32_t)chunks.size();
        for (int i = off[s]; i < off[s + 1]; ++i) sys_of[i] = s;
        if (sys_theta) th[s] = sys_theta[s];
        for (int g = 0; g < sys_n[s]; g += 64) groups.push_back(make_int4(off[s] + g, std::min(64, sys_n[s] - g), s, 0));
        samp[s + 1] = samp[s] + (sys_n[s] + nbx::kSubtreeMax - 1) / nbx::kSubtreeMax;
    }
    b->ngroups = (int)groups.size();
    b->nchunk = (int)chunks.size();
    const size_t N = total, M = m64, S = nsys;
    const size_t tiles = (N + nbx::kSortTile - 1) / nbx::kSortTile;
    bool ok = dalloc(&b->st, 1) && dalloc(&b->off, S + 1) && dalloc(&b->sys_of, N) && dalloc(&b->theta, S) &&
              dalloc(&b->groups, groups.size()) && dalloc(&b->ext, 4 * S) && dalloc(&b->mm, 6 * S) &&
              dalloc(&b->keys, N) && dalloc(&b->skeys, N) && dalloc(&b->kbuf, N) && dalloc(&b->k2, N) &&
              dalloc(&b->k2buf, N) && dalloc(&b->perm, N) && dalloc(&b->perm1, N) && dalloc(&b->vbuf, N) &&
              dalloc(&b->pos, N) && dalloc(&b->sorted, N) && dalloc(&b->hist, tiles * nbx::kRadix + nbx::kRadix) &&
              dalloc(&b->child, 8 * M) && dalloc(&b->index, 1) && dalloc(&b->count, M) && dalloc(&b->start, M) &&
              dalloc(&b->cell_level, M) &&
              dalloc(&b->sum64, M) && dalloc(&b->nodes4, M) &&
              dalloc(&b->fnode, M + 8) && dalloc(&b->fmeta, M + 8) && dalloc(&b->boff, M) &&
              dalloc(&b->bsum, M / nbx::kOffRun + 2) && dalloc(&b->froot4, S) &&
              dalloc(&b->frootm, S) && dalloc(&b->pblk, M) && dalloc(&b->bpar, N) && dalloc(&b->counts, N) && dalloc(&b->acc_scratch, 3 * N) && dalloc(&b->bodies4, N) &&
              dalloc(&b->walk, 2 * groups.size()) && dalloc(&b->gcost, groups.size()) &&
              dalloc(&b->gkey, N) && dalloc(&b->gkey2, N) && dalloc(&b->gorder, N) &&
              dalloc(&b->samp_off, S + 1) && dalloc(&b->bchunk, chunks.size()) && dalloc(&b->bchunk_off, S + 1) &&
              dalloc(&b->bpart, 6 * chunks.size()) && nbx::alloc_tree_scratch(b->ts, (int)N, nsys, samp[nsys]);
    for (int k = 0; ok && k < 8; ++k) ok = hipEventCreate(&b->stage_ev[k]) == hipSuccess;
    ok = ok && hipMemset(b->st, 0, sizeof(nbx::DeviceState)) == hipSuccess &&
         hipMemset(b->gcost, 0, sizeof(int32_t) * groups.size()) == hipSuccess &&
         hipMemcpy(b->off, off.data(), 4 * (S + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->sys_of, sys_of.data(), 4 * N, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->theta, th.data(), 4 * S, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->groups, groups.data(), sizeof(int4) * groups.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->samp_off, samp.data(), 4 * (S + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->bchunk, chunks.data(), sizeof(int4) * chunks.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->bchunk_off, choff.data(), 4 * (S + 1), hipMemcpyHostToDevice) == hipSuccess &&
         hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        nbx_batch_destroy(b);
        return bfail(nullptr, NBX_ERR_HIP, "device allocation failed");
    }
    *out = b;
    return NBX_OK;
}

int nbx_batch_destroy(nbx_batch *b)
{
    if (!b) return NBX_OK;
    hipSetDevice(b->device);
    hipDeviceSynchronize();
    drop_graph(b);
    void *bufs[] = {b->st, b->off, b->sys_of, b->theta, b->groups, b->ext, b->mm, b->keys, b->skeys, b->kbuf,
                    b->k2, b->k2buf, b->perm, b->perm1, b->vbuf, b->pos, b->sorted, b->hist, b->child, b->index,
                    b->count, b->start, b->cell_level, b->sum64, b->nodes4,
                    b->fnode, b->fmeta, b->boff, b->bsum, b->froot4, b->frootm, b->pblk, b->bpar, b->counts, b->acc_scratch, b->samp_off, b->bodies4, b->gkey, b->gkey2, b->gorder, b->walk, b->gcost,
                    b->bchunk, b->bchunk_off, b->bpart};
    for (void *q : bufs)
        if (q) hipFree(q);
    nbx::free_tree_scratch(b->ts);
    for (auto &ev : b->stage_ev)
        if (ev) hipEventDestroy(ev);
    delete b;
    return NBX_OK;
}

const char *nbx_batch_last_error(const nbx_batch *b) { return b ? b->last_error.c_str() : g_batch_error.c_str(); }

int nbx_batch_bind(nbx_batch *b, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax,
                   float *ay, float *az, float *mass)
{
    if (!b) return NBX_ERR_INVALID;
    if (!x || !y || !z || !vx || !vy || !vz || !ax || !ay || !az || !mass) return bfail(b, NBX_ERR_INVALID, "null array");
    b->x = x; b->y = y; b->z = z; b->vx = vx; b->vy = vy; b->vz = vz;
    b->ax = ax; b->ay = ay; b->az = az; b->mass = mass;
    drop_graph(b);
    return NBX_OK;
}

int nbx_batch_step(nbx_batch *b, void *stream, float *stage_ms)
{
    if (!b) return NBX_ERR_INVALID;
    if (!bound(b)) return bfail(b, NBX_ERR_ORDER, "nbx_batch_bind first");
    hipSetDevice(b->device);
    hipStream_t s = (hipStream_t)stream;
    enqueue_step(b, s, stage_ms ? b->stage_ev : nullptr);
    if (int r = finish(b)) return r;
    if (stage_ms) {
        if (hipEventSynchronize(b->stage_ev[7]) != hipSuccess) return bfail(b, NBX_ERR_HIP, "hipEventSynchronize");
        for (int k = 0; k < 7; ++k) hipEventElapsedTime(&stage_ms[k], b->stage_ev[k], b->stage_ev[k + 1]);
    }
    return NBX_OK;
}

int nbx_batch_graph_step(nbx_batch *b, void *stream)
{
    if (!b) return NBX_ERR_INVALID;
    if (!bound(b)) return bfail(b, NBX_ERR_ORDER, "nbx_batch_bind first");
    if (!stream) return bfail(b, NBX_ERR_INVALID, "graph capture needs a non-null stream");
    hipSetDevice(b->device);
    hipStream_t s = (hipStream_t)stream;
    if (!b->exec) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return bfail(b, NBX_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        enqueue_step(b, s, nullptr);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || !g) return bfail(b, NBX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            return bfail(b, NBX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    hipError_t e = hipGraphLaunch(b->exec, s);
    if (e != hipSuccess) return bfail(b, NBX_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return NBX_OK;
}

int nbx_batch_check(nbx_batch *b, void *stream)
{
    if (!b) return NBX_ERR_INVALID;
    hipSetDevice(b->device);
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return bfail(b, NBX_ERR_HIP, "hipStreamSynchronize");
    // err, sticky and seen as in nbx_check: reported once, then sticky cleared
    // and this step's bits marked seen
    int e[3] = {0, 0, 0};
    static_assert(offsetof(nbx::DeviceState, sticky) == offsetof(nbx::DeviceState, err) + sizeof(int) &&
                      offsetof(nbx::DeviceState, seen) == offsetof(nbx::DeviceState, err) + 2 * sizeof(int),
                  "err, sticky and seen are adjacent");
    if (hipMemcpy(e, &b->st->err, 3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(b, NBX_ERR_HIP, "hipMemcpy");
    if (e[1] || e[2] != e[0]) {
        const int upd[2] = {0, e[0]};
        if (hipMemcpy(&b->st->sticky, upd, 2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return bfail(b, NBX_ERR_HIP, "hipMemcpy");
    }
    const int err = e[0] | e[1];
    if (err & nbx::kErrNonFinite) return bfail(b, NBX_ERR_NONFINITE, "non-finite position or zero cell mass");
    if (err & nbx::kErrCapacity) return bfail(b, NBX_ERR_CAPACITY, "tree needs more cells than the ensemble holds");
    if (err & nbx::kErrBucket) return bfail(b, NBX_ERR_BUCKET, "more than 8 bodies share one key");
    if (err & nbx::kErrStack) return bfail(b, NBX_ERR_STACK, "traversal stack overflow");
    if (err & nbx::kErrStall) return bfail(b, NBX_ERR_HIP, "top summary: its workgroups did not run together");
    return NBX_OK;
}

int nbx_batch_count_interactions(nbx_batch *b, int64_t *counts, void *stream)
{
    if (!b || !counts) return NBX_ERR_INVALID;
    hipSetDevice(b->device);
    nbx::launch_batch_force(b, (hipStream_t)stream, counts, b->force_waves);
    return finish(b);
}

int nbx_batch_walk_stats(nbx_batch *b, int64_t *out, void *stream)
{
    if (!b || !out) return NBX_ERR_INVALID;
    hipSetDevice(b->device);
    hipStream_t s = (hipStream_t)stream;
    nbx::launch_batch_force(b, s, b->counts, b->force_waves, b->walk);
    if (int r = finish(b)) return r;
    if (hipStreamSynchronize(s) != hipSuccess) return bfail(b, NBX_ERR_HIP, "hipStreamSynchronize");
    std::vector<int64_t> cnt(b->N), wk(2 * (size_t)b->ngroups);
    if (hipMemcpy(cnt.data(), b->counts, 8 * (size_t)b->N, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(wk.data(), b->walk, 16 * (size_t)b->ngroups, hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(b, NBX_ERR_HIP, "copy failed");
    int64_t inter = 0, steps = 0, frames = 0;
    for (int64_t v : cnt) inter += v;
    for (int k = 0; k < b->ngroups; ++k) {
        steps += wk[2 * (size_t)k];
        frames += wk[2 * (size_t)k + 1];
    }
    out[0] = inter;
    out[1] = steps;
    out[2] = frames;
    out[3] = b->ngroups;
    return NBX_OK;
}

int nbx_batch_autotune(nbx_batch *b, void *stream, int *waves_out)
{
    if (!b) return NBX_ERR_INVALID;
    if (!bound(b)) return bfail(b, NBX_ERR_ORDER, "nbx_batch_bind first");
    hipSetDevice(b->device);
    hipStream_t s = (hipStream_t)stream;
    // tree of the current state, without the update
    nbx::launch_batch_reset(b, s);
    nbx::launch_batch_bbox(b, s);
    nbx::launch_batch_keys(b, s);
    nbx::launch_batch_sort(b, s);
    nbx::launch_batch_build(b, s);
    nbx::launch_batch_summarize(b, s);
    nbx::launch_batch_layout(b, s);
    const int cand[3] = {2, 4, 8};
    float best = 1e30f;
    int pick = b->force_waves;
    // time into scratch so that the bound accelerations are left untouched
    float *ax = b->ax, *ay = b->ay, *az = b->az;
    b->ax = b->acc_scratch;
    b->ay = b->acc_scratch + b->N;
    b->az = b->acc_scratch + 2 * (size_t)b->N;
    for (int w : cand) {
        nbx::launch_batch_force(b, s, nullptr, w);  // warm-up
        hipEventRecord(b->stage_ev[0], s);
        for (int r = 0; r < 3; ++r) nbx::launch_batch_force(b, s, nullptr, w);
        hipEventRecord(b->stage_ev[1], s);
        if (hipEventSynchronize(b->stage_ev[1]) != hipSuccess) {
            b->ax = ax; b->ay = ay; b->az = az;
            return bfail(b, NBX_ERR_HIP, "hipEventSynchronize");
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, b->stage_ev[0], b->stage_ev[1]);
        if (ms < best) { best = ms; pick = w; }
    }
    b->ax = ax; b->ay = ay; b->az = az;
    b->force_waves = pick;
    drop_graph(b);
    if (waves_out) *waves_out = pick;
    return finish(b);
}

int nbx_batch_copy_buffer(nbx_batch *b, int which, void *dst, int64_t count, void *stream)
{
    if (!b || !dst || count < 0) return NBX_ERR_INVALID;
    hipSetDevice(b->device);
    const void *src = nullptr;
    size_t elem = 0, cap = 0;
    switch (which) {
    case NBX_BUF_KEYS: src = b->keys; elem = 8; cap = b->N; break;
    case NBX_BUF_SORTED_KEYS: src = b->skeys; elem = 8; cap = b->N; break;
    case NBX_BUF_PERM: src = b->perm; elem = 4; cap = b->N; break;
    case NBX_BUF_EXTENT: src = b->ext; elem = 8; cap = 4 * (size_t)b->S; break;
    case NBX_BUF_CELL_LEVEL: src = b->cell_level; elem = 4; cap = b->m; break;
    case NBX_BUF_CELL_MASS64: elem = 8; cap = b->m; break;
    case NBX_BUF_CELL_COM64: elem = 8; cap = 3 * (size_t)b->m; break;
    default: return bfail(b, NBX_ERR_INVALID, "unknown buffer");
    }
    if ((size_t)count > cap) return bfail(b, NBX_ERR_INVALID, "count exceeds buffer");
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return bfail(b, NBX_ERR_HIP, "copy failed");
    const hipError_t e = which == NBX_BUF_CELL_MASS64 || which == NBX_BUF_CELL_COM64
                             ? nbx::copy_cell_sums(b->sum64, which == NBX_BUF_CELL_COM64, static_cast<double *>(dst),
                                                   (size_t)count)
                             : hipMemcpy(dst, src, elem * (size_t)count, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bfail(b, NBX_ERR_HIP, "copy failed");
    return NBX_OK;
}

// Single-system autotuner: the current tree must be laid out (after nbx_sort).
int nbx_autotune(nbx_context *c, void *stream, int *waves_out)
{
    if (!c) return NBX_ERR_INVALID;
    if (c->stage < 4) {
        c->last_error = "nbx_autotune needs nbx_sort first";
        return NBX_ERR_ORDER;
    }
    hipSetDevice(c->device);
    hipStream_t s = (hipStream_t)stream;
    const int cand[3] = {2, 4, 8};
    float best = 1e30f;
    int pick = c->force_waves;
    float *ax = c->acc_scratch, *ay = ax + c->n, *az = ay + c->n;
    for (int w : cand) {
        nbx::launch_force_waves(c, s, ax, ay, az, 1.f, w);
        hipEventRecord(c->ev0, s);
        for (int r = 0; r < 3; ++r) nbx::launch_force_waves(c, s, ax, ay, az, 1.f, w);
        hipEventRecord(c->ev1, s);
        if (hipEventSynchronize(c->ev1) != hipSuccess) {
            c->last_error = "hipEventSynchronize failed";
            return NBX_ERR_HIP;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        if (ms < best) { best = ms; pick = w; }
    }
    c->force_waves = pick;
    if (waves_out) *waves_out = pick;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->last_error = std::string("kernel launch: ") + hipGetErrorString(e);
        return NBX_ERR_HIP;
    }
    return NBX_OK;
}

}  // extern "C"
