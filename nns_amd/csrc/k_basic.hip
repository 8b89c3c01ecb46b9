// k_basic.hip -- streaming stages: reset, bounding box, keys, update.
// All of them are HBM-bound: coalesced grid-stride loops, no LDS except the
// bounding-box block reduction and the Hilbert state table of the keys.
#include "nbx_device.h"

namespace nbx {

// resetArraysKernel semantics (3D/CUDA/CUDA_NBody/src/Kernels.cu:21-64).  The
// eight child slots of a node are written as two 16-byte stores.
__global__ void __launch_bounds__(256) reset_kernel(
    int *__restrict__ mutex, float *__restrict__ x, float *__restrict__ y, float *__restrict__ z,
    float *__restrict__ mass, int *__restrict__ count, int *__restrict__ start, int *__restrict__ sorted,
    int *__restrict__ child, int *__restrict__ index, float *minX, float *maxX, float *minY, float *maxY,
    float *minZ, float *maxZ, int n, int m, DeviceState *st)
{
    const int stride = blockDim.x * gridDim.x;
    const int4 neg = make_int4(-1, -1, -1, -1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        int4 *c4 = reinterpret_cast<int4 *>(child + 8 * (int64_t)i);
        c4[0] = neg;
        c4[1] = neg;
        if (i < n) {
            count[i] = 1;
        } else {
            x[i] = 0.f; y[i] = 0.f; z[i] = 0.f; mass[i] = 0.f;
            count[i] = 0;
        }
        start[i] = -1;
        sorted[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *mutex = 0;
        *index = n;
        *minX = 0.f; *maxX = 0.f; *minY = 0.f; *maxY = 0.f; *minZ = 0.f; *maxZ = 0.f;
        st->sticky |= st->err & ~st->seen;  // an unreported error of this step waits for nbx_check
        st->seen = 0;
        st->err = 0;
    }
}

__device__ inline float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Bounding box in two launches: every workgroup folds its grid-stride share of
// the bodies into six partial extremes, part[6 b .. 6 b + 6); one workgroup
// then folds the partials and derives the FP64 extent of the key grid: origin
// (xMin, yMin, zMin) and maxSpan = max |max - min| over the axes
// (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:143-147).  The launch boundary
// makes the partials visible to the fold.  (A last-workgroup ticket in one
// launch needs a device-scope release from every workgroup, and on this GPU
// each such release writes back the XCD's L2.)

// min x, max x, min y, max y, min z, max z of the workgroup, in v on thread 0.
__device__ inline void block_extremes(float (&v)[6], float (*red)[4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; k += 2) {
        v[k] = wave_min(v[k]);
        v[k + 1] = wave_max(v[k + 1]);
    }
    if (lane == 0)
        for (int k = 0; k < 6; ++k) red[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            for (int k = 0; k < 6; k += 2) {
                v[k] = fminf(v[k], red[k][w]);
                v[k + 1] = fmaxf(v[k + 1], red[k + 1][w]);
            }
}

// Folds bodies i = first, first + stride, ... < end into v (min x, max x, min y,
// max y, min z, max z); true if a coordinate is not finite.  Four strides per
// iteration, their twelve loads issued before any is used (min and max are
// exact, so the order of the folds does not matter).
__device__ inline bool fold_extremes(const float *__restrict__ x, const float *__restrict__ y,
                                     const float *__restrict__ z, int first, int end, int stride, float (&v)[6])
{
    bool bad = false;
    auto fold = [&](float xi, float yi, float zi) {
        bad |= !(isfinite(xi) && isfinite(yi) && isfinite(zi));
        v[0] = fminf(v[0], xi); v[1] = fmaxf(v[1], xi);
        v[2] = fminf(v[2], yi); v[3] = fmaxf(v[3], yi);
        v[4] = fminf(v[4], zi); v[5] = fmaxf(v[5], zi);
    };
    int i = first;
    for (; i < end - 3 * stride; i += 4 * stride) {
        float xs[4], ys[4], zs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            xs[u] = x[i + u * stride];
            ys[u] = y[i + u * stride];
            zs[u] = z[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) fold(xs[u], ys[u], zs[u]);
    }
    for (; i < end; i += stride) fold(x[i], y[i], z[i]);
    return bad;
}

__global__ void __launch_bounds__(256) bbox_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                   const float *__restrict__ z, int n, float *__restrict__ part,
                                                   DeviceState *st)
{
    __shared__ float red[6][4];
    const int tid = threadIdx.x;
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    const bool bad = fold_extremes(x, y, z, blockIdx.x * blockDim.x + tid, n, blockDim.x * gridDim.x, v);
    if (__ballot(bad) && (tid & 63) == 0) atomicOr(&st->err, kErrNonFinite);
    block_extremes(v, red);
    if (tid == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = v[k];
}

__global__ void __launch_bounds__(256) bbox_fold_kernel(const float *__restrict__ part, int nparts, DeviceState *st,
                                                        float *minX, float *maxX, float *minY, float *maxY,
                                                        float *minZ, float *maxZ)
{
    __shared__ float red[6][4];
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        const float *q = part + 6 * b;
        for (int k = 0; k < 6; k += 2) {
            v[k] = fminf(v[k], q[k]);
            v[k + 1] = fmaxf(v[k + 1], q[k + 1]);
        }
    }
    block_extremes(v, red);
    if (threadIdx.x == 0) {
        *minX = v[0]; *maxX = v[1]; *minY = v[2];
        *maxY = v[3]; *minZ = v[4]; *maxZ = v[5];
        double xs = fabs((double)v[1] - (double)v[0]);
        double ys = fabs((double)v[3] - (double)v[2]);
        double zs = fabs((double)v[5] - (double)v[4]);
        double sp = xs > ys ? xs : ys;
        sp = sp < zs ? zs : sp;
        st->ext[0] = v[0]; st->ext[1] = v[2]; st->ext[2] = v[4]; st->ext[3] = sp;
    }
}

// Keys, and (x, y, z, m) packed as float4 for the one-line gather of the layout.
// Hilbert keys map the Morton key through the state table in LDS; Morton keys
// need neither the table nor the barrier after its fill.
template <bool HILBERT>
__global__ void __launch_bounds__(256) keys_kernel(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const float *__restrict__ mass, int n, const DeviceState *__restrict__ st,
    uint64_t *__restrict__ keys, float4 *__restrict__ bodies4)
{
    __shared__ uint8_t s_tab[HILBERT ? kHilbertTable : 1];
    if (HILBERT) {
        hilbert_table_fill(s_tab);
        __syncthreads();
    }
    const double ext[4] = {st->ext[0], st->ext[1], st->ext[2], st->ext[3]};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const float xi = x[i], yi = y[i], zi = z[i];
        uint64_t k = morton_key(xi, yi, zi, ext);
        if (HILBERT) k = hilbert_top_lds(k & ~(1ull << 63), kKeyLevels, s_tab);
        keys[i] = k;
        if (mass) bodies4[i] = make_float4(xi, yi, zi, mass[i]);
    }
}

// updateKernel (Kernels.cu:581-600): v += dt a; x += d dt v, with explicit fused
// multiply-adds so that the result does not depend on compiler contraction.
// After an error of this step (latch raised) the bodies are left where they
// are: the accelerations were not computed.
__global__ void __launch_bounds__(1024) update_kernel(
    float *__restrict__ x, float *__restrict__ y, float *__restrict__ z, float *__restrict__ vx,
    float *__restrict__ vy, float *__restrict__ vz, const float *__restrict__ ax,
    const float *__restrict__ ay, const float *__restrict__ az, int n, float dt, float d, const DeviceState *st)
{
    if (st && latched(st)) return;
    const float ddt = d * dt;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        float u = fmaf(dt, ax[i], vx[i]);
        float v = fmaf(dt, ay[i], vy[i]);
        float w = fmaf(dt, az[i], vz[i]);
        vx[i] = u; vy[i] = v; vz[i] = w;
        x[i] = fmaf(ddt, u, x[i]);
        y[i] = fmaf(ddt, v, y[i]);
        z[i] = fmaf(ddt, w, z[i]);
    }
}

// ------------------------------------------------------------ stars ("suns")
// Stars are bodies outside the tree (3D/OpenMP/OpenMP_NBody/src/
// Interaction.cpp:111-124): each star pulls every body and every other star
// directly and is pulled back by every body, with the Plummer kernel
// m d (|d|^2 + eps^2)^-3/2; a zero distance contributes nothing (:28-30).
// The terms are evaluated in FP64.  Bodies: one thread each, the sum over the
// stars added once to the FP32 acceleration the tree walk wrote.  Stars: each
// block reduces its bodies' pulls per star (wave shuffles, then LDS) to one
// partial per (block, star); star_final_kernel adds the partials in block
// order and the star-star terms: no atomics, the same sums on every run.
constexpr int kStarThreads = 256;

__device__ inline double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ void __launch_bounds__(kStarThreads) star_bodies_kernel(
    const float *__restrict__ sx, const float *__restrict__ sy, const float *__restrict__ sz,
    const float *__restrict__ sm, int ns, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ z, const float *__restrict__ mass, float *__restrict__ ax, float *__restrict__ ay,
    float *__restrict__ az, int n, double G, double eps2, double *__restrict__ part, const DeviceState *st)
{
    __shared__ double s_red[kStarThreads / 64][NBX_MAX_STARS][3];
    if (latched(st)) return;  // one value for the whole launch: nothing in it raises the latch
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double qx[NBX_MAX_STARS], qy[NBX_MAX_STARS], qz[NBX_MAX_STARS], qm[NBX_MAX_STARS];
    double px[NBX_MAX_STARS], py[NBX_MAX_STARS], pz[NBX_MAX_STARS];  // this thread's pulls on each star
#pragma unroll
    for (int s = 0; s < NBX_MAX_STARS; ++s) {
        const bool on = s < ns;
        qx[s] = on ? sx[s] : 0.0; qy[s] = on ? sy[s] : 0.0; qz[s] = on ? sz[s] : 0.0; qm[s] = on ? sm[s] : 0.0;
        px[s] = py[s] = pz[s] = 0.0;
    }
    for (int i = blockIdx.x * blockDim.x + tid; i < n; i += blockDim.x * gridDim.x) {
        const double bx = x[i], by = y[i], bz = z[i], bm = mass[i];
        double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int s = 0; s < NBX_MAX_STARS; ++s) {
            if (s >= ns) break;
            const double dx = qx[s] - bx, dy = qy[s] - by, dz = qz[s] - bz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0) continue;
            const double r2 = d2 + eps2;
            const double w = 1.0 / (r2 * sqrt(r2));
            gx += qm[s] * w * dx; gy += qm[s] * w * dy; gz += qm[s] * w * dz;
            px[s] -= bm * w * dx; py[s] -= bm * w * dy; pz[s] -= bm * w * dz;
        }
        ax[i] += (float)(G * gx);
        ay[i] += (float)(G * gy);
        az[i] += (float)(G * gz);
    }
#pragma unroll
    for (int s = 0; s < NBX_MAX_STARS; ++s) {
        if (s >= ns) break;
        const double a = wave_sum_f64(px[s]), b = wave_sum_f64(py[s]), c = wave_sum_f64(pz[s]);
        if (lane == 0) { s_red[wave][s][0] = a; s_red[wave][s][1] = b; s_red[wave][s][2] = c; }
    }
    __syncthreads();
    for (int k = tid; k < 3 * ns; k += blockDim.x) {
        const int s = k / 3, c = k - 3 * s;
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += s_red[w][s][c];
        part[((int64_t)blockIdx.x * NBX_MAX_STARS + s) * 3 + c] = v;
    }
}

__global__ void __launch_bounds__(64) star_final_kernel(
    const float *__restrict__ sx, const float *__restrict__ sy, const float *__restrict__ sz,
    const float *__restrict__ sm, int ns, float *__restrict__ sax, float *__restrict__ say, float *__restrict__ saz,
    double G, double eps2, const double *__restrict__ part, int nblocks, const DeviceState *st)
{
    if (latched(st)) return;
    const int s = threadIdx.x;
    if (s >= ns) return;
    const double qx = sx[s], qy = sy[s], qz = sz[s];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int t = 0; t < ns; ++t) {  // Interaction.cpp:113-118: every other star
        if (t == s) continue;
        const double dx = (double)sx[t] - qx, dy = (double)sy[t] - qy, dz = (double)sz[t] - qz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 == 0.0) continue;
        const double r2 = d2 + eps2;
        const double w = (double)sm[t] / (r2 * sqrt(r2));
        gx += w * dx; gy += w * dy; gz += w * dz;
    }
    for (int b = 0; b < nblocks; ++b) {  // :120-123: the bodies, block partials in order
        const double *p = part + ((int64_t)b * NBX_MAX_STARS + s) * 3;
        gx += p[0]; gy += p[1]; gz += p[2];
    }
    sax[s] = (float)(G * gx);
    say[s] = (float)(G * gy);
    saz[s] = (float)(G * gz);
}

__global__ void fill_f32_kernel(float *p, float v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)blockDim.x * gridDim.x)
        p[i] = v;
}

static inline int grid_for(int64_t work, int block, int cap)
{
    int64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void launch_reset(nbx_context *c, hipStream_t s, int *mutex, float *x, float *y, float *z, float *mass,
                  int *count, int *start, int *sorted, int *child, int *index, float *minX, float *maxX,
                  float *minY, float *maxY, float *minZ, float *maxZ, int n, int m)
{
    const int block = 256;
    reset_kernel<<<grid_for(m, block, 8192), block, 0, s>>>(mutex, x, y, z, mass, count, start, sorted, child,
                                                            index, minX, maxX, minY, maxY, minZ, maxZ, n, m,
                                                            c->st);
}

void launch_bbox(nbx_context *c, hipStream_t s, const float *x, const float *y, const float *z,
                 float *minX, float *maxX, float *minY, float *maxY, float *minZ, float *maxZ, int n)
{
    const int block = 256;
    int g = grid_for(n, block * 8, c->bbox_blocks);
    bbox_kernel<<<g, block, 0, s>>>(x, y, z, n, c->bbox_part, c->st);
    bbox_fold_kernel<<<1, block, 0, s>>>(c->bbox_part, g, c->st, minX, maxX, minY, maxY, minZ, maxZ);
}

void launch_keys(nbx_context *c, hipStream_t s, const float *x, const float *y, const float *z, const float *mass,
                 int n, int key_type, uint64_t *keys)
{
    const int block = 256;
    if (key_type == NBX_KEY_HILBERT)
        keys_kernel<true><<<grid_for(n, block, 65536), block, 0, s>>>(x, y, z, mass, n, c->st, keys, c->bodies4);
    else
        keys_kernel<false><<<grid_for(n, block, 65536), block, 0, s>>>(x, y, z, mass, n, c->st, keys, c->bodies4);
}

void launch_update(hipStream_t s, int block, const DeviceState *st, float *x, float *y, float *z, float *vx, float *vy,
                   float *vz, const float *ax, const float *ay, const float *az, int n, float dt, float d)
{
    if (block <= 0) block = 256;
    update_kernel<<<grid_for(n, block, 65536), block, 0, s>>>(x, y, z, vx, vy, vz, ax, ay, az, n, dt, d, st);
}

void launch_fill_f32(hipStream_t s, float *p, float v, int64_t n)
{
    if (n <= 0) return;
    fill_f32_kernel<<<grid_for(n, 256, 4096), 256, 0, s>>>(p, v, n);
}

void launch_star_forces(nbx_context *c, hipStream_t s, const float *sx, const float *sy, const float *sz,
                        const float *sm, float *sax, float *say, float *saz, int ns, const float *x, const float *y,
                        const float *z, const float *mass, float *ax, float *ay, float *az, int n, float g)
{
    const double eps2 = (double)layout_eps2(c->p);
    const int blocks = n > 0 ? grid_for(n, kStarThreads, kStarPartBlocks) : 1;
    star_bodies_kernel<<<blocks, kStarThreads, 0, s>>>(sx, sy, sz, sm, ns, x, y, z, mass, ax, ay, az, n, (double)g,
                                                       eps2, c->star_part, c->st);
    star_final_kernel<<<1, 64, 0, s>>>(sx, sy, sz, sm, ns, sax, say, saz, (double)g, eps2, c->star_part, blocks,
                                       c->st);
}

// ------------------------------------------------------------ batched form
// Root cell N + s of every system: empty slots, level 0.
__global__ void __launch_bounds__(256) reset_batch_kernel(int S, int N, int *__restrict__ child, int *index,
                                                          int32_t *__restrict__ cell_level, DeviceState *st)
{
    const int4 neg = make_int4(-1, -1, -1, -1);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < S; s += blockDim.x * gridDim.x) {
        const int64_t c = (int64_t)N + s;
        int4 *c4 = reinterpret_cast<int4 *>(child + 8 * c);
        c4[0] = neg;
        c4[1] = neg;
        cell_level[c] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *index = N + S;
        st->sticky |= st->err & ~st->seen;
        st->seen = 0;
        st->err = 0;
    }
}

// Ensemble: one workgroup per chunk of at most kBboxChunk bodies of one system
// (chunk[c] = {system, first, end}) folds its partial extremes into part[6 c ..];
// then one thread per system folds its chunks [choff[s], choff[s+1]) into
// mm[6 s ..] and the FP64 extent ext[4 s ..].  A workgroup per system would
// leave the largest system's loop setting the launch's time.
__global__ void __launch_bounds__(256) bbox_batch_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                         const float *__restrict__ z, const int4 *__restrict__ chunk,
                                                         float *__restrict__ part, DeviceState *st)
{
    __shared__ float red[6][4];
    const int tid = threadIdx.x;
    const int4 ck = chunk[blockIdx.x];
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    const bool bad = fold_extremes(x, y, z, ck.y + tid, ck.z, blockDim.x, v);
    if (__ballot(bad) && (tid & 63) == 0) atomicOr(&st->err, kErrNonFinite);
    block_extremes(v, red);
    if (tid == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = v[k];
}

__global__ void __launch_bounds__(256) bbox_batch_fold_kernel(int S, const int32_t *__restrict__ choff,
                                                              const float *__restrict__ part, float *__restrict__ mm,
                                                              double *__restrict__ ext)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (int c = choff[s]; c < choff[s + 1]; ++c)
        for (int k = 0; k < 6; k += 2) {
            v[k] = fminf(v[k], part[6 * c + k]);
            v[k + 1] = fmaxf(v[k + 1], part[6 * c + k + 1]);
        }
    for (int k = 0; k < 6; ++k) mm[6 * s + k] = v[k];
    double xs = fabs((double)v[1] - (double)v[0]);
    double ys = fabs((double)v[3] - (double)v[2]);
    double zs = fabs((double)v[5] - (double)v[4]);
    double sp = xs > ys ? xs : ys;
    sp = sp < zs ? zs : sp;
    ext[4 * s + 0] = v[0]; ext[4 * s + 1] = v[2]; ext[4 * s + 2] = v[4]; ext[4 * s + 3] = sp;
}

__global__ void __launch_bounds__(256) keys_batch_kernel(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const float *__restrict__ mass, int n, const int32_t *__restrict__ sys_of, const double *__restrict__ ext,
    uint64_t *__restrict__ keys, float4 *__restrict__ bodies4)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const int s = sys_of[i];
        const double e[4] = {ext[4 * s], ext[4 * s + 1], ext[4 * s + 2], ext[4 * s + 3]};
        const float xi = x[i], yi = y[i], zi = z[i];
        keys[i] = morton_key(xi, yi, zi, e);
        bodies4[i] = make_float4(xi, yi, zi, mass[i]);
    }
}

void launch_batch_reset(nbx_batch *b, hipStream_t s)
{
    reset_batch_kernel<<<grid_for(b->S, 256, 1024), 256, 0, s>>>(b->S, b->N, b->child, b->index, b->cell_level,
                                                                 b->st);
}

void launch_batch_bbox(nbx_batch *b, hipStream_t s)
{
    bbox_batch_kernel<<<b->nchunk, 256, 0, s>>>(b->x, b->y, b->z, b->bchunk, b->bpart, b->st);
    bbox_batch_fold_kernel<<<(b->S + 255) / 256, 256, 0, s>>>(b->S, b->bchunk_off, b->bpart, b->mm, b->ext);
}

void launch_batch_keys(nbx_batch *b, hipStream_t s)
{
    keys_batch_kernel<<<grid_for(b->N, 256, 65536), 256, 0, s>>>(b->x, b->y, b->z, b->mass, b->N, b->sys_of,
                                                                 b->ext, b->keys, b->bodies4);
}

}  // namespace nbx
