// k_adjoint.hip -- vector-Jacobian product of the Barnes-Hut accelerations.
//
// Given u_i = dL/da_i for every body, the walk of the force kernel is repeated
// on the same tree and frame layout (so each body meets exactly the nodes of
// its forward interaction list) and, for every interaction of body i with a
// node n of mass M at X (d = X - x_i, s = (|d|^2 + eps^2)^-3/2),
//     h      = M s (u - 3 (u.d) d / (|d|^2 + eps^2))      = d(u.a_i)/dX
//     dL/dx_i -= h,   dL/dX_n += h,   dL/dM_n += s (u.d).
// A zero distance contributes nothing (the forward skips it).  The node terms
// of one child step are summed over the lanes that took it, through LDS: each
// step writes one float4 per lane (rows padded to 65 entries, so the
// transposed reads below are free of bank conflicts), and every kVjpBatch = 4
// recorded steps the wave reads them back transposed (16 lanes per step, 4
// entries each), folds the 16 partial sums with four DPP moves and adds one
// FP64 atomic per component and node.  Bodies are nodes too: their sums go to
// the body's own entry.  LDS stays under 7 KB per wave, so 16 waves fit a CU.
// The sums have one copy per XCD (the wave's XCC_ID hardware register picks
// it): a line that only one XCD updates stays in that XCD's L2.  The copy is a
// placement choice, not a correctness one -- every copy takes device-scope
// atomics -- and the flow-down pass adds the copies up first.
//
// A cell's X and M are sums over its bodies, X = sum m_k x_k / M, so the node
// gradients flow down: for every body k
//     dL/dx_k += m_k A,   dL/dm_k += B + A.x_k,
// with A = sum of G_c / M_c and B = sum of (GM_c - G_c.X_c / M_c) over the
// cells c that contain k.  The walk attributes a collapsed one-child chain to
// its bottom cell; the cells of a chain hold the same bodies, mass and centre,
// so the sums over the ancestors are unchanged.
#include "nbx_device.h"

namespace nbx {

constexpr int kVjpBatch = 4;
constexpr int kXcdCopies = 8;

// Sum over each aligned group of 16 lanes (a DPP row), in every lane of the row.
__device__ inline float seg16_sum(float v)
{
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
    return v;
}

__device__ inline int xcd_copy()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return (int)(v & (kXcdCopies - 1));
}

template <int W>
__global__ void __launch_bounds__(64 * W) vjp_kernel(
    int n, int ngroups, const float4 *__restrict__ nodes4, const float4 *__restrict__ fnode,
    const int2 *__restrict__ fmeta, const float4 *__restrict__ froot4, const int2 *__restrict__ frootm,
    const int32_t *__restrict__ perm, const DeviceState *st, float eps2, float G,
    const float *__restrict__ ux, const float *__restrict__ uy, const float *__restrict__ uz,
    float *__restrict__ gx, float *__restrict__ gy, float *__restrict__ gz, double *__restrict__ tgt,
    int64_t copy_stride, int *__restrict__ err)
{
    __shared__ uint4 stk[W][kForceStack];
    __shared__ float4 red[W][kVjpBatch][65];
    __shared__ int rid[W][kVjpBatch];
    if (latched(st)) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = xcd_swizzle(blockIdx.x, gridDim.x) * W + wave;
    if (g >= ngroups) return;

    const int first = g * 64;
    const int cnt_bodies = min(64, n - first);
    double *const tcopy = tgt + copy_stride * xcd_copy();

    const int p = first + lane;
    const bool valid = lane < cnt_bodies;
    const float4 me = valid ? nodes4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int i = valid ? perm[p] : 0;
    const float vx = valid ? G * ux[i] : 0.f;
    const float vy = valid ? G * uy[i] : 0.f;
    const float vz = valid ? G * uz[i] : 0.f;
    const uint64_t live = __ballot(valid);

    float ox = 0.f, oy = 0.f, oz = 0.f;  // dL/dx of the lane's own body
    int nrec = 0;
    float4(*rw)[65] = red[wave];
    int *ri = rid[wave];

    // Lane 16s + q sums entries 4q .. 4q+3 of recorded step s; lane 16s adds the total.
    auto flush = [&](const int count) {
        __builtin_amdgcn_wave_barrier();
        const int s_ = lane >> 4, q_ = lane & 15;
        float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
        if (s_ < count) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float4 v = rw[s_][4 * q_ + l];
                ax += v.x; ay += v.y; az += v.z; aw += v.w;
            }
        }
        ax = seg16_sum(ax);
        ay = seg16_sum(ay);
        az = seg16_sum(az);
        aw = seg16_sum(aw);
        if (q_ == 0 && s_ < count) {
            double *t = tcopy + 4 * (int64_t)ri[s_];
            unsafeAtomicAdd(t + 0, (double)ax);
            unsafeAtomicAdd(t + 1, (double)ay);
            unsafeAtomicAdd(t + 2, (double)az);
            unsafeAtomicAdd(t + 3, (double)aw);
        }
        __builtin_amdgcn_wave_barrier();
    };

    // One child step (descriptor {tag, crit} as in the force walk).
    auto step = [&](const float4 q, const int tag, const float crit, const uint64_t mask) -> uint64_t {
        const float dx = q.x - me.x;
        const float dy = q.y - me.y;
        const float dz = q.z - me.z;
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const uint64_t ok = __ballot(crit < d2);
        const uint64_t omask = mask & ~ok;
        const uint64_t takem = mask & ok & __ballot(d2 > 0.f);
        if (takem) {
            const bool sel = __builtin_amdgcn_inverse_ballot_w64(takem);
            const float r2 = sel ? d2 + eps2 : 1.f;  // finite on every lane
            const float rs = __builtin_amdgcn_rsqf(r2);
            const float rs2 = rs * rs;
            const float s3 = sel ? rs * rs2 : 0.f;
            const float ud = fmaf(vz, dz, fmaf(vy, dy, vx * dx));
            const float k3 = 3.f * ud * rs2;
            const float ms = q.w * s3;
            const float hx = ms * fmaf(-k3, dx, vx);
            const float hy = ms * fmaf(-k3, dy, vy);
            const float hz = ms * fmaf(-k3, dz, vz);
            const float hw = s3 * ud;
            ox -= hx;
            oy -= hy;
            oz -= hz;
            rw[nrec][lane] = make_float4(hx, hy, hz, hw);
            if (lane == 0) ri[nrec] = (tag & 15) == 0 ? (tag >> 4) : n + (tag >> 4);
            nrec = __builtin_amdgcn_readfirstlane(nrec + 1);
            if (nrec == kVjpBatch) {
                flush(kVjpBatch);
                nrec = 0;
            }
        }
        return omask;
    };

    uint4 *s = stk[wave];
    int sp = 0;
    {
        const int2 rd = frootm[0];
        const int rt = __builtin_amdgcn_readfirstlane(rd.x);
        const float rc = __int_as_float(__builtin_amdgcn_readfirstlane(rd.y));
        const uint64_t om = step(froot4[0], rt, rc, live);
        if (om) {
            if (lane == 0) s[0] = make_uint4((unsigned)rt, 0u, (unsigned)om, (unsigned)(om >> 32));
            sp = 1;
        }
    }
    while (sp > 0) {
        if (sp > kForceStack - 7) {
            if (lane == 0) atomicOr(err, kErrStack);
            break;
        }
        sp = __builtin_amdgcn_readfirstlane(sp - 1);
        const uint4 e = s[sp];
        const int tag = __builtin_amdgcn_readfirstlane((int)e.x);
        const int nch = tag & 15;
        const int blk = tag >> 4;
        const uint64_t mask = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.w) << 32) |
                              (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.z);
        const float4 *bq = fnode + 8 * (int64_t)blk;
        const int2 *bd = fmeta + 8 * (int64_t)blk;
        float4 q[8];
        int2 d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            q[j] = bq[j];
            d[j] = bd[j];
        }
        asm volatile("" ::"s"(q[0].x), "s"(q[0].y), "s"(q[0].z), "s"(q[0].w), "s"(q[1].x), "s"(q[1].y),
                     "s"(q[1].z), "s"(q[1].w), "s"(q[2].x), "s"(q[2].y), "s"(q[2].z), "s"(q[2].w), "s"(q[3].x),
                     "s"(q[3].y), "s"(q[3].z), "s"(q[3].w));
        asm volatile("" ::"s"(q[4].x), "s"(q[4].y), "s"(q[4].z), "s"(q[4].w), "s"(q[5].x), "s"(q[5].y),
                     "s"(q[5].z), "s"(q[5].w), "s"(q[6].x), "s"(q[6].y), "s"(q[6].z), "s"(q[6].w), "s"(q[7].x),
                     "s"(q[7].y), "s"(q[7].z), "s"(q[7].w));
        asm volatile("" ::"s"(d[0].x), "s"(d[0].y), "s"(d[1].x), "s"(d[1].y), "s"(d[2].x), "s"(d[2].y),
                     "s"(d[3].x), "s"(d[3].y), "s"(d[4].x), "s"(d[4].y), "s"(d[5].x), "s"(d[5].y), "s"(d[6].x),
                     "s"(d[6].y), "s"(d[7].x), "s"(d[7].y));
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j >= nch) break;
            const int tj = __builtin_amdgcn_readfirstlane(d[j].x);
            const float cj = __int_as_float(__builtin_amdgcn_readfirstlane(d[j].y));
            const uint64_t om = step(q[j], tj, cj, mask);
            if (om) {
                if (lane == 0) s[sp] = make_uint4((unsigned)tj, 0u, (unsigned)om, (unsigned)(om >> 32));
                sp = __builtin_amdgcn_readfirstlane(sp + 1);
            }
        }
    }
    if (nrec > 0) flush(nrec);
    if (valid) {
        gx[i] = ox;
        gy[i] = oy;
        gz[i] = oz;
    }
}

__global__ void __launch_bounds__(256) clear_f64_kernel(double *__restrict__ p, int64_t count)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)blockDim.x * gridDim.x)
        p[k] = 0.0;
}

// Per block b: A and B summed over b and its ancestors (walk up pblk).
// Adds the per-XCD copies into copy 0 (count values each).
__global__ void __launch_bounds__(256) adj_fold_kernel(double *__restrict__ tgt, int64_t count, int64_t copy_stride,
                                                       const DeviceState *st)
{
    if (latched(st)) return;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)blockDim.x * gridDim.x) {
        double v = tgt[k];
        for (int c = 1; c < kXcdCopies; ++c) v += tgt[copy_stride * c + k];
        tgt[k] = v;
    }
}

__global__ void __launch_bounds__(256) adj_cells_kernel(int n, const int *__restrict__ index,
                                                        const int32_t *__restrict__ pblk,
                                                        const double *__restrict__ tgt,
                                                        const double *__restrict__ mass64,
                                                        const double *__restrict__ mom64, double *__restrict__ ab,
                                                        const DeviceState *st)
{
    if (latched(st)) return;
    const int nblk = *index - n + 1;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += blockDim.x * gridDim.x) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, bb = 0.0;
        for (int c = b; c >= 0; c = pblk[c]) {
            const double *t = tgt + 4 * ((int64_t)n + c);
            double M, X, Y, Z;
            if (c == 0) {
                M = st->root64[0]; X = st->root64[1]; Y = st->root64[2]; Z = st->root64[3];
            } else {
                const int64_t cell = (int64_t)c + n - 1;
                M = mass64[cell];
                X = mom64[3 * cell] / M; Y = mom64[3 * cell + 1] / M; Z = mom64[3 * cell + 2] / M;
            }
            if (M > 0.0) {
                const double g0 = t[0] / M, g1 = t[1] / M, g2 = t[2] / M;
                a0 += g0; a1 += g1; a2 += g2;
                bb += t[3] - (g0 * X + g1 * Y + g2 * Z);
            } else {
                bb += t[3];
            }
        }
        double *o = ab + 4 * (int64_t)b;
        o[0] = a0; o[1] = a1; o[2] = a2; o[3] = bb;
    }
}

// Per body (sorted position p): the node terms of the body itself plus the
// flow-down of the cells above it.  gx/gy/gz hold the walk's own part.
__global__ void __launch_bounds__(256) adj_bodies_kernel(int n, const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ bpar,
                                                         const float4 *__restrict__ nodes4,
                                                         const double *__restrict__ tgt,
                                                         const double *__restrict__ ab, float *__restrict__ gx,
                                                         float *__restrict__ gy, float *__restrict__ gz,
                                                         float *__restrict__ gm, const DeviceState *st)
{
    if (latched(st)) return;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += blockDim.x * gridDim.x) {
        const int i = perm[p];
        const float4 b = nodes4[p];
        const double *t = tgt + 4 * (int64_t)p;
        const double *a = ab + 4 * (int64_t)bpar[p];
        const double m = b.w;
        gx[i] = (float)((double)gx[i] + t[0] + m * a[0]);
        gy[i] = (float)((double)gy[i] + t[1] + m * a[1]);
        gz[i] = (float)((double)gz[i] + t[2] + m * a[2]);
        gm[i] = (float)(t[3] + a[3] + a[0] * (double)b.x + a[1] * (double)b.y + a[2] * (double)b.z);
    }
}

int vjp_copies() { return kXcdCopies; }

void launch_vjp(nbx_context *c, hipStream_t s, const float *ux, const float *uy, const float *uz, float *gx,
                float *gy, float *gz, float *gm, float g)
{
    const int n = c->n;
    if (n <= 0) return;
    if (c->layout_theta != c->p.theta) launch_frame_layout(c, s);
    const int64_t ntgt = 4 * ((int64_t)n + (c->m - n + 1));  // bodies, then blocks
    const int64_t stride = 4 * ((int64_t)c->m_max + 1);      // one copy
    for (int k = 0; k < kXcdCopies; ++k)
        clear_f64_kernel<<<(int)std::min<int64_t>((ntgt + 255) / 256, 65536), 256, 0, s>>>(c->adj_tgt + stride * k,
                                                                                           ntgt);
    constexpr int W = 2;
    const int ngroups = (n + 63) / 64;
    vjp_kernel<W><<<(ngroups + W - 1) / W, 64 * W, 0, s>>>(
        n, ngroups, c->nodes4, c->fnode, c->fmeta, c->froot4, c->frootm, c->perm, c->st,
        c->p.softening * c->p.softening, g, ux, uy, uz, gx, gy, gz, c->adj_tgt, stride, &c->st->err);
    adj_fold_kernel<<<(int)std::min<int64_t>((ntgt + 255) / 256, 65536), 256, 0, s>>>(c->adj_tgt, ntgt, stride,
                                                                                      c->st);
    const int64_t nblk = c->m - n + 1;
    adj_cells_kernel<<<(int)std::min<int64_t>((nblk + 255) / 256, 65536), 256, 0, s>>>(
        n, c->index, c->pblk, c->adj_tgt, c->mass64, c->mom64, c->adj_ab, c->st);
    adj_bodies_kernel<<<(int)std::min<int64_t>((n + 255) / 256, 65536), 256, 0, s>>>(
        n, c->perm, c->bpar, c->nodes4, c->adj_tgt, c->adj_ab, gx, gy, gz, gm, c->st);
}

}  // namespace nbx
