// k_adjoint.hip -- vector-Jacobian product of the Barnes-Hut accelerations.
//
// Given u_i = dL/da_i for every body, the walk of the force kernel is repeated
// on the same tree and frame layout (so each body meets exactly the nodes of
// its forward interaction list) and, for every interaction of body i with a
// node n of mass M at X (d = X - x_i, s = (|d|^2 + eps^2)^-3/2),
//     h      = M s (u - 3 (u.d) d / (|d|^2 + eps^2))      = d(u.a_i)/dX
//     dL/dx_i -= h,   dL/dX_n += h,   dL/dM_n += s (u.d).
// The lane's own body contributes nothing (its forward term vanishes
// identically); another body at the same point does when eps > 0 (h = M eps^-3
// u) and is skipped, as in the forward walk, when eps = 0.  The node terms
// of one child step are summed over the lanes that took it, through LDS: each
// step writes one float4 per lane (rows padded to 65 entries, so the
// transposed reads below are free of bank conflicts), and every kVjpBatch = 4
// recorded steps the wave reads them back transposed (16 lanes per step, 4
// entries each) and folds the 16 partials with four DPP moves.  LDS stays under
// 7 KB per wave, so 16 waves fit a CU.  Bodies are nodes too.
//
// The per-step sums do not go to the nodes by atomics (one FP64 atomic per
// component and step is a memory-side transaction: at N = 262K they cost more
// than the walk).  They are appended to a log {32-bit node id, float4}: each
// wave takes chunks of kChunk records with one atomicAdd and seals the unused
// tail of a chunk with a sentinel node id.  The log is radix-sorted by node id
// (stable; 3 passes while the ids -- sorted positions of bodies, n + the first
// frame record of a cell -- stay below 2^24 - 1) and summed per node in FP64, in
// log order, over tiles of 64 sorted records (one per lane); only the node runs
// that cross a tile take an atomic.  The log order is that of the chunk grants,
// so a node's FP64 total can differ from run to run in its last bits.  A wave
// that finds the log full adds its remaining sums with atomics directly.
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
constexpr int kChunk = 256;   // log records taken per atomicAdd

// Sum over each aligned group of 16 lanes (a DPP row), in every lane of the row.
__device__ inline float seg16_sum(float v)
{
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
    return v;
}

template <int W, bool SOFT>
__global__ void __launch_bounds__(64 * W) vjp_kernel(
    int n, int ngroups, const float4 *__restrict__ nodes4, const float4 *__restrict__ fnode,
    const int2 *__restrict__ fmeta, const float4 *__restrict__ froot4, const int2 *__restrict__ frootm,
    const int32_t *__restrict__ perm, const int32_t *__restrict__ gorder, const DeviceState *st, float eps2,
    float G,
    const float *__restrict__ ux, const float *__restrict__ uy, const float *__restrict__ uz,
    float *__restrict__ gx, float *__restrict__ gy, float *__restrict__ gz, double *__restrict__ tgt,
    uint32_t *__restrict__ log_id, float4 *__restrict__ log_val, int64_t cap, uint32_t sentinel, DeviceState *dst,
    int *__restrict__ err)
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

    const bool valid = lane < cnt_bodies;
    const int p = valid ? gorder[first + lane] : 0;  // the lane's body, by sorted position
    const float4 me = valid ? nodes4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int i = valid ? perm[p] : 0;
    const int me_tag = valid ? p << 4 : -1;  // the descriptor tag of the lane's own body
    const float vx = valid ? G * ux[i] : 0.f;
    const float vy = valid ? G * uy[i] : 0.f;
    const float vz = valid ? G * uz[i] : 0.f;
    const uint64_t live = __ballot(valid);

    float ox = 0.f, oy = 0.f, oz = 0.f;  // dL/dx of the lane's own body
    int nrec = 0;
    int64_t cpos = 0, cend = 0;  // the wave's current log chunk [cpos, cend)
    bool full = false;           // the log is full: add sums with atomics
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
        if (!full && cend - cpos < count) {
            if (lane < cend - cpos) log_id[cpos + lane] = sentinel;  // seal the old chunk (< 4 records)
            int b = 0;
            if (lane == 0) b = atomicAdd(&dst->log_next, kChunk);
            b = __builtin_amdgcn_readfirstlane(b);
            if ((int64_t)b + kChunk > cap) {
                full = true;
            } else {
                cpos = b;
                cend = (int64_t)b + kChunk;
            }
        }
        if (q_ == 0 && s_ < count) {
            if (!full) {
                log_id[cpos + s_] = (uint32_t)ri[s_];
                log_val[cpos + s_] = make_float4(ax, ay, az, aw);
            } else {
                double *t = tgt + 4 * (int64_t)ri[s_];
                unsafeAtomicAdd(t + 0, (double)ax);
                unsafeAtomicAdd(t + 1, (double)ay);
                unsafeAtomicAdd(t + 2, (double)az);
                unsafeAtomicAdd(t + 3, (double)aw);
            }
        }
        if (!full) cpos += count;
        __builtin_amdgcn_wave_barrier();
    };

    // One child step (descriptor {tag, crit} as in the force walk).
    auto step = [&](const float4 q, const int tag, const float crit, const uint64_t mask) -> uint64_t {
        const float dx = q.x - me.x;
        const float dy = q.y - me.y;
        const float dz = q.z - me.z;
        const float d2e = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));  // |d|^2 + eps^2
        const uint64_t ok = __ballot(crit < d2e);  // the force walk's test
        const uint64_t omask = mask & ~ok;
        // The lane's own term vanishes identically (a body's tag is its sorted
        // position << 4); another body at the same point keeps its derivative
        // when eps > 0, and is left out, as in the forward walk, when eps = 0.
        const uint64_t keep = SOFT ? ~__ballot(tag == me_tag) : __ballot(d2e > 0.f);
        const uint64_t takem = mask & ok & keep;
        if (takem) {
            const bool sel = __builtin_amdgcn_inverse_ballot_w64(takem);
            const float r2 = SOFT ? d2e : (sel ? d2e : 1.f);  // finite on every lane (d2e >= eps^2 > 0 when SOFT)
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
    bool overflow = false;
    while (sp > 0) {  // one exit: the accumulators keep their registers between frames
        if (sp > kForceStack - 7) {
            overflow = true;
            sp = 0;
            continue;
        }
        sp = __builtin_amdgcn_readfirstlane(sp - 1);
        const uint4 e = s[sp];
        const int tag = __builtin_amdgcn_readfirstlane((int)e.x);
        const int nch = tag & 15;
        const int off = tag >> 4;  // the block's first record
        const uint64_t mask = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.w) << 32) |
                              (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.z);
        const float4 *bq = fnode + off;
        const int2 *bd = fmeta + off;
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
    if (overflow && lane == 0) atomicOr(err, kErrStack);
    if (nrec > 0) flush(nrec);
    if (!full)
        for (int64_t k = cpos + lane; k < cend; k += 64) log_id[k] = sentinel;  // seal the last chunk
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
// Sort length: every granted chunk is written in full (records or sentinels).
__global__ void log_count_kernel(DeviceState *st, int64_t cap)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t granted = (cap / kChunk) * kChunk;
        st->log_count = (int)min((int64_t)st->log_next, granted);
    }
}

// Per-node sums of the sorted log in tiles of 64 records (one per lane); a wave
// takes kReduceTiles consecutive tiles and issues all their loads (ids, then the
// values gathered from the log) before the first sum.  Per tile: an inclusive
// scan over the lanes holding the same node (six shuffle steps), then the last
// lane of each run adds it to the node.  A run that continues into the previous
// or the next tile is added with an FP64 atomic; a run inside the tile with a
// plain add (no other tile holds that node).
constexpr int kReduceTiles = 4;

__global__ void __launch_bounds__(256) log_reduce_kernel(const uint32_t *__restrict__ sk,
                                                         const int32_t *__restrict__ sv,
                                                         const float4 *__restrict__ val, const DeviceState *st,
                                                         uint32_t sentinel, double *__restrict__ tgt)
{
    if (latched(st)) return;
    const int64_t count = st->log_count;
    const int lane = threadIdx.x & 63;
    const int64_t ntiles = (count + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 6) * kReduceTiles;
    for (int64_t w0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kReduceTiles; w0 < ntiles;
         w0 += wstride) {
        uint32_t ids[kReduceTiles];
        int32_t src[kReduceTiles];
        float4 vals[kReduceTiles];
#pragma unroll
        for (int t = 0; t < kReduceTiles; ++t) {
            const int64_t k = (w0 + t) * 64 + lane;
            ids[t] = k < count ? sk[k] : sentinel;
            src[t] = k < count ? sv[k] : 0;
        }
#pragma unroll
        for (int t = 0; t < kReduceTiles; ++t)
            vals[t] = ids[t] != sentinel ? val[src[t]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int t = 0; t < kReduceTiles; ++t) {
            if (w0 + t >= ntiles) break;
            const int64_t k0 = (w0 + t) * 64, k = k0 + lane;
            const uint32_t id = ids[t];
            const bool live = id != sentinel;
            float4 v = vals[t];
            const int id32 = live ? (int)id : -1;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int oid = __shfl_up(id32, o);
                const float tx = __shfl_up(v.x, o), ty = __shfl_up(v.y, o), tz = __shfl_up(v.z, o),
                            tw = __shfl_up(v.w, o);
                if (lane >= o && oid == id32) {
                    v.x += tx; v.y += ty; v.z += tz; v.w += tw;
                }
            }
            const int next = __shfl_down(id32, 1);
            const bool last_of_run = live && (lane == 63 || next != id32);
            const int lane0_id = __shfl(id32, 0);
            const bool from_prev = k0 > 0 && lane0_id == id32 && sk[k0 - 1] == id;
            const bool into_next = lane == 63 && k + 1 < count && sk[k + 1] == id;
            if (last_of_run) {
                double *p = tgt + 4 * (int64_t)id;
                if (from_prev || into_next) {
                    unsafeAtomicAdd(p + 0, (double)v.x); unsafeAtomicAdd(p + 1, (double)v.y);
                    unsafeAtomicAdd(p + 2, (double)v.z); unsafeAtomicAdd(p + 3, (double)v.w);
                } else {
                    p[0] += v.x; p[1] += v.y; p[2] += v.z; p[3] += v.w;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256) adj_cells_kernel(int n, const int *__restrict__ index,
                                                        const int32_t *__restrict__ pblk,
                                                        const uint32_t *__restrict__ boff,
                                                        const uint32_t *__restrict__ bsum,
                                                        const double *__restrict__ tgt,
                                                        const double4 *__restrict__ sum64, double *__restrict__ ab,
                                                        const DeviceState *st)
{
    if (latched(st)) return;
    const int nblk = *index - n + 1;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += blockDim.x * gridDim.x) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, bb = 0.0;
        for (int c = b; c >= 0; c = pblk[c]) {
            const double *t = tgt + 4 * ((int64_t)n + frame_off(boff, bsum, c));  // the block's node id
            double M, X, Y, Z;
            if (c == 0) {
                M = st->root64[0]; X = st->root64[1]; Y = st->root64[2]; Z = st->root64[3];
            } else {
                const int64_t cell = (int64_t)c + n - 1;
                const double4 cs = sum64[cell];
                M = cs.x;
                X = cs.y / M; Y = cs.z / M; Z = cs.w / M;
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

// Radix passes over the node ids: a body's id is its sorted position, a cell's
// n + the first record of its frame block, so ids are below n_max + m_max and
// the sentinel, 2^(kRadixBits passes) - 1, must exceed them.
static int sort_passes(int64_t id_max)
{
    int p = 1;
    while ((int64_t(1) << (kRadixBits * p)) < id_max + 2) ++p;
    return p;
}

__global__ void vjp_reset_kernel(DeviceState *st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) st->log_next = 0;
}

template <class T> static bool dmalloc(T **p, int64_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (size_t)(count > 0 ? count : 1)) == hipSuccess;
}

bool alloc_vjp_scratch(nbx_context *c)
{
    const int64_t cnt = 4 * ((int64_t)c->n_max + c->m_max + 8);  // node ids: bodies, then n + a record
    const int64_t ngroups = ((int64_t)c->n_max + 63) / 64;
    // about 60 records per body at theta = 0.5; a full log falls back to atomics
    int64_t cap = 64 * (int64_t)c->n_max + (int64_t)kChunk * (ngroups + 1);
    cap = std::min<int64_t>(cap, INT32_MAX - kSortTile);
    const bool even = sort_passes((int64_t)c->n_max + c->m_max) % 2 == 0;
    const int64_t tiles = (cap + kSortTile - 1) / kSortTile;
    bool ok = dmalloc(&c->adj_tgt, cnt) && dmalloc(&c->adj_ab, cnt) && dmalloc(&c->adj_log_id, cap) &&
              dmalloc(&c->adj_log_val, cap) && dmalloc(&c->adj_sk, cap) && dmalloc(&c->adj_sv, cap) &&
              dmalloc(&c->adj_tv, cap) && (!even || dmalloc(&c->adj_tk, cap)) &&
              dmalloc(&c->adj_hist, tiles * kRadix + kRadix);
    if (!ok) {
        free_vjp_scratch(c);
        return false;
    }
    c->adj_cap = cap;
    return true;
}

void free_vjp_scratch(nbx_context *c)
{
    void *p[] = {c->adj_tgt, c->adj_ab, c->adj_log_id, c->adj_log_val, c->adj_sk, c->adj_sv, c->adj_tv, c->adj_tk,
                 c->adj_hist};
    for (void *q : p)
        if (q) hipFree(q);
    c->adj_tgt = c->adj_ab = nullptr;
    c->adj_log_id = c->adj_sk = c->adj_tk = nullptr;
    c->adj_log_val = nullptr;
    c->adj_sv = c->adj_tv = nullptr;
    c->adj_hist = nullptr;
    c->adj_cap = 0;
}

void launch_vjp(nbx_context *c, hipStream_t s, const float *ux, const float *uy, const float *uz, float *gx,
                float *gy, float *gz, float *gm, float g)
{
    const int n = c->n;
    if (n <= 0) return;
    if (frame_layout_stale(c)) launch_frame_layout(c, s);
    const int64_t ntgt = 4 * ((int64_t)n + c->m + 8);  // bodies, then cells by their block's first record
    clear_f64_kernel<<<(int)std::min<int64_t>((ntgt + 255) / 256, 65536), 256, 0, s>>>(c->adj_tgt, ntgt);
    vjp_reset_kernel<<<1, 1, 0, s>>>(c->st);
    const int passes = sort_passes((int64_t)c->n_max + c->m_max);
    const uint32_t sentinel = (uint32_t)((uint64_t(1) << (kRadixBits * passes)) - 1);
    constexpr int W = 2;
    const int ngroups = (n + 63) / 64;
    const float eps2 = layout_eps2(c->p);
    auto walk = eps2 > 0.f ? vjp_kernel<W, true> : vjp_kernel<W, false>;
    walk<<<(ngroups + W - 1) / W, 64 * W, 0, s>>>(
        n, ngroups, c->nodes4, c->fnode, c->fmeta, c->froot4, c->frootm, c->perm, c->gorder, c->st, eps2, g, ux, uy,
        uz, gx, gy, gz, c->adj_tgt, c->adj_log_id, c->adj_log_val, c->adj_cap, sentinel, c->st, &c->st->err);
    log_count_kernel<<<1, 1, 0, s>>>(c->st, c->adj_cap);
    uint32_t *tk = passes % 2 ? c->adj_log_id : c->adj_tk;  // the input keys are dead after the first pass
    launch_radix_passes(c->adj_hist, s, c->adj_log_id, nullptr, (int)c->adj_cap, 0, passes, c->adj_sk, c->adj_sv, tk,
                        c->adj_tv, nullptr, &c->st->log_count);
    const int64_t waves = (c->adj_cap + 64 * kReduceTiles - 1) / (64 * kReduceTiles);
    log_reduce_kernel<<<(int)std::min<int64_t>((waves + 3) / 4, 65536), 256, 0, s>>>(
        c->adj_sk, c->adj_sv, c->adj_log_val, c->st, sentinel, c->adj_tgt);
    const int64_t nblk = c->m - n + 1;
    adj_cells_kernel<<<(int)std::min<int64_t>((nblk + 255) / 256, 65536), 256, 0, s>>>(
        n, c->index, c->pblk, c->boff, c->bsum, c->adj_tgt, c->sum64, c->adj_ab, c->st);
    adj_bodies_kernel<<<(int)std::min<int64_t>((n + 255) / 256, 65536), 256, 0, s>>>(
        n, c->perm, c->bpar, c->nodes4, c->adj_tgt, c->adj_ab, gx, gy, gz, gm, c->st);
}

}  // namespace nbx
