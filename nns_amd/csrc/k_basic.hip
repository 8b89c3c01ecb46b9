// k_basic.hip -- streaming stages: reset, bounding box, keys, update.
// All of them are HBM-bound: coalesced grid-stride loops, no LDS except the
// bounding-box block reduction.
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
        st->bbox_ticket = 0u;
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

// Bounding box: per-block min/max, then the last block to finish (ticket
// counter) folds the partials and derives the FP64 extent of the key grid:
// origin (xMin, yMin, zMin) and maxSpan = max |max - min| over the axes
// (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:143-147).
__global__ void __launch_bounds__(256) bbox_kernel(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z, int n,
    float *__restrict__ part, DeviceState *st, float *minX, float *maxX, float *minY, float *maxY,
    float *minZ, float *maxZ)
{
    __shared__ float red[6][4];
    __shared__ int s_last, s_bad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY, e = INFINITY, f = -INFINITY;
    int bad = 0;
    for (int i = blockIdx.x * blockDim.x + tid; i < n; i += blockDim.x * gridDim.x) {
        float xi = x[i], yi = y[i], zi = z[i];
        bad |= !(isfinite(xi) && isfinite(yi) && isfinite(zi));
        a = fminf(a, xi); b = fmaxf(b, xi);
        c = fminf(c, yi); d = fmaxf(d, yi);
        e = fminf(e, zi); f = fmaxf(f, zi);
    }
    if (tid == 0) s_bad = 0;
    a = wave_min(a); b = wave_max(b); c = wave_min(c); d = wave_max(d); e = wave_min(e); f = wave_max(f);
    __syncthreads();
    if (bad) s_bad = 1;
    if (lane == 0) {
        red[0][wave] = a; red[1][wave] = b; red[2][wave] = c;
        red[3][wave] = d; red[4][wave] = e; red[5][wave] = f;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 1; w < nw; ++w) {
            red[0][0] = fminf(red[0][0], red[0][w]); red[1][0] = fmaxf(red[1][0], red[1][w]);
            red[2][0] = fminf(red[2][0], red[2][w]); red[3][0] = fmaxf(red[3][0], red[3][w]);
            red[4][0] = fminf(red[4][0], red[4][w]); red[5][0] = fmaxf(red[5][0], red[5][w]);
        }
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = red[k][0];
        if (s_bad) atomicOr(&st->err, kErrNonFinite);
        __threadfence();
        unsigned t = atomicAdd(&st->bbox_ticket, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // last block: fold all partials
    a = INFINITY; b = -INFINITY; c = INFINITY; d = -INFINITY; e = INFINITY; f = -INFINITY;
    for (int k = tid; k < (int)gridDim.x; k += blockDim.x) {
        const float *q = part + 6 * k;
        a = fminf(a, __hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        b = fmaxf(b, __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        c = fminf(c, __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        d = fmaxf(d, __hip_atomic_load(q + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        e = fminf(e, __hip_atomic_load(q + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        f = fmaxf(f, __hip_atomic_load(q + 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    a = wave_min(a); b = wave_max(b); c = wave_min(c); d = wave_max(d); e = wave_min(e); f = wave_max(f);
    __syncthreads();
    if (lane == 0) {
        red[0][wave] = a; red[1][wave] = b; red[2][wave] = c;
        red[3][wave] = d; red[4][wave] = e; red[5][wave] = f;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 1; w < nw; ++w) {
            red[0][0] = fminf(red[0][0], red[0][w]); red[1][0] = fmaxf(red[1][0], red[1][w]);
            red[2][0] = fminf(red[2][0], red[2][w]); red[3][0] = fmaxf(red[3][0], red[3][w]);
            red[4][0] = fminf(red[4][0], red[4][w]); red[5][0] = fmaxf(red[5][0], red[5][w]);
        }
        *minX = red[0][0]; *maxX = red[1][0]; *minY = red[2][0];
        *maxY = red[3][0]; *minZ = red[4][0]; *maxZ = red[5][0];
        double xs = fabs((double)red[1][0] - (double)red[0][0]);
        double ys = fabs((double)red[3][0] - (double)red[2][0]);
        double zs = fabs((double)red[5][0] - (double)red[4][0]);
        double sp = xs > ys ? xs : ys;
        sp = sp < zs ? zs : sp;
        st->ext[0] = red[0][0]; st->ext[1] = red[2][0]; st->ext[2] = red[4][0]; st->ext[3] = sp;
        st->bbox_ticket = 0u;
    }
}

// Keys, and (x, y, z, m) packed as float4 for the one-line gather of the layout.
__global__ void __launch_bounds__(256) keys_kernel(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const float *__restrict__ mass, int n, const DeviceState *__restrict__ st, int key_type,
    uint64_t *__restrict__ keys, float4 *__restrict__ bodies4)
{
    const double ext[4] = {st->ext[0], st->ext[1], st->ext[2], st->ext[3]};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const float xi = x[i], yi = y[i], zi = z[i];
        uint64_t k = morton_key(xi, yi, zi, ext);
        if (key_type == NBX_KEY_HILBERT) k = hilbert_of(k & ~(1ull << 63));
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
    bbox_kernel<<<g, block, 0, s>>>(x, y, z, n, c->bbox_part, c->st, minX, maxX, minY, maxY, minZ, maxZ);
}

void launch_keys(nbx_context *c, hipStream_t s, const float *x, const float *y, const float *z, const float *mass,
                 int n, int key_type, uint64_t *keys)
{
    const int block = 256;
    keys_kernel<<<grid_for(n, block, 65536), block, 0, s>>>(x, y, z, mass, n, c->st, key_type, keys, c->bodies4);
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

// One block per system: min/max over [off[s], off[s+1]) and the FP64 extent.
__global__ void __launch_bounds__(256) bbox_batch_kernel(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const int32_t *__restrict__ off, float *__restrict__ mm, double *__restrict__ ext, DeviceState *st)
{
    __shared__ float red[6][4];
    const int s = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lo = off[s], hi = off[s + 1];
    float a = INFINITY, b = -INFINITY, c = INFINITY, d = -INFINITY, e = INFINITY, f = -INFINITY;
    int bad = 0;
    for (int i = lo + tid; i < hi; i += blockDim.x) {
        float xi = x[i], yi = y[i], zi = z[i];
        bad |= !(isfinite(xi) && isfinite(yi) && isfinite(zi));
        a = fminf(a, xi); b = fmaxf(b, xi);
        c = fminf(c, yi); d = fmaxf(d, yi);
        e = fminf(e, zi); f = fmaxf(f, zi);
    }
    if (__ballot(bad) && lane == 0) atomicOr(&st->err, kErrNonFinite);
    a = wave_min(a); b = wave_max(b); c = wave_min(c); d = wave_max(d); e = wave_min(e); f = wave_max(f);
    if (lane == 0) {
        red[0][wave] = a; red[1][wave] = b; red[2][wave] = c;
        red[3][wave] = d; red[4][wave] = e; red[5][wave] = f;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = blockDim.x >> 6;
        for (int w = 1; w < nw; ++w) {
            red[0][0] = fminf(red[0][0], red[0][w]); red[1][0] = fmaxf(red[1][0], red[1][w]);
            red[2][0] = fminf(red[2][0], red[2][w]); red[3][0] = fmaxf(red[3][0], red[3][w]);
            red[4][0] = fminf(red[4][0], red[4][w]); red[5][0] = fmaxf(red[5][0], red[5][w]);
        }
        for (int k = 0; k < 6; ++k) mm[6 * s + k] = red[k][0];
        double xs = fabs((double)red[1][0] - (double)red[0][0]);
        double ys = fabs((double)red[3][0] - (double)red[2][0]);
        double zs = fabs((double)red[5][0] - (double)red[4][0]);
        double sp = xs > ys ? xs : ys;
        sp = sp < zs ? zs : sp;
        ext[4 * s + 0] = red[0][0]; ext[4 * s + 1] = red[2][0]; ext[4 * s + 2] = red[4][0]; ext[4 * s + 3] = sp;
    }
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
    bbox_batch_kernel<<<b->S, 256, 0, s>>>(b->x, b->y, b->z, b->off, b->mm, b->ext, b->st);
}

void launch_batch_keys(nbx_batch *b, hipStream_t s)
{
    keys_batch_kernel<<<grid_for(b->N, 256, 65536), 256, 0, s>>>(b->x, b->y, b->z, b->mass, b->N, b->sys_of,
                                                                 b->ext, b->keys, b->bodies4);
}

}  // namespace nbx
