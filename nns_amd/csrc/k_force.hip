// k_force.hip -- Barnes-Hut accelerations: 64-lane wavefront traversal.
//
// One wave owns a group of up to 64 consecutive bodies in key order, one body
// per lane.  The wave walks the tree depth first with an LDS stack of entries
// {node, level, lane mask}.  The node of an iteration is the same for every
// lane: it is read once from LDS, made scalar with readfirstlane, and its
// float4 comes in through the scalar cache (s_load_dwordx4).  Each lane applies
// its own opening test to a cell (l/D < theta written as crit_L < |d|^2 with
// crit_L = (l_L/theta)^2; the reference CPU test of
// 3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:64-67).  The test is a ballot, so
// the lane masks stay in SGPRs: lanes that accept add the cell's monopole under
// exec = mask & accept, lanes that reject vote the cell open and its children
// are pushed with mask & ~accept.  Every lane therefore sees exactly its own
// per-particle interaction list; the wave pays for the union of the 64 lists.
// Leaves (bodies) are always direct (Interaction.cpp:60-63); a zero distance
// contributes nothing (:28-30).
//
// Interaction: a += m d (|d|^2 + eps^2)^-3/2, d = q - x (the 20-flop form of
// GPU Gems 3 ch. 31), with the raw v_rsq_f32 (|d|^2 + eps^2 is never denormal
// when eps > 0; the unsoftened variant masks d = 0).  With Kahan on, monopoles
// go into a plain partial that is folded into a compensated total every 16
// nodes.
//
// Blocks are remapped so that the blocks one XCD receives (dispatch is
// round-robin over the 8 XCDs) walk consecutive groups and share its L2.
//
// Single system: group g is sorted positions [64g, 64g+64), the root is the
// special entry kRootRef (data in DeviceState, slots at tchild[0..7]).
// Ensemble: group g comes from a table {first, count, system}; the walk starts
// at the system's root cell with the system's own theta and extent.
#include "nbx_device.h"

namespace nbx {

template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH, int W>
__global__ void __launch_bounds__(64 * W) force_kernel(
    int n, int ngroups, const float4 *__restrict__ nodes4, const int32_t *__restrict__ tchild,
    const int32_t *__restrict__ perm, const DeviceState *st, float theta, float eps2, float G,
    float *__restrict__ ax, float *__restrict__ ay, float *__restrict__ az, int64_t *__restrict__ counts,
    int *__restrict__ err, const int4 *__restrict__ groups, const float *__restrict__ sys_theta,
    const double *__restrict__ sys_ext)
{
    __shared__ uint4 stk[W][kForceStack];
    if (latched(st)) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = xcd_swizzle(blockIdx.x, gridDim.x) * W + wave;
    if (g >= ngroups) return;

    int first, cnt_bodies, root_ref;
    double span;
    float th;
    if (BATCH) {
        const int4 gr = groups[g];
        first = gr.x;
        cnt_bodies = gr.y;
        root_ref = n + gr.z;
        span = sys_ext[4 * gr.z + 3];
        th = sys_theta[gr.z];
    } else {
        first = g * 64;
        cnt_bodies = min(64, n - first);
        root_ref = kRootRef;
        span = st->ext[3];
        th = theta;
    }
    // crit_0 = (l_0/theta)^2, l_0 = side of the root cell of the key grid
    const double lc0 = ldexp(span / kMaskCoord, kKeyLevels);
    const double r0 = lc0 / (double)th;
    const float crit0 = (float)(r0 * r0);
    float4 root = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!BATCH) root = st->root4;

    const int p = first + lane;
    const bool valid = lane < cnt_bodies;
    const float4 me = valid ? nodes4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t live = __ballot(valid);

    float sx = 0.f, sy = 0.f, sz = 0.f;  // compensated totals
    float cx = 0.f, cy = 0.f, cz = 0.f;  // compensations
    float px = 0.f, py = 0.f, pz = 0.f;  // partials of the current tile
    int64_t cnt = 0;
    int tile = 0;

    uint4 *s = stk[wave];
    if (lane == 0) s[0] = make_uint4((unsigned)root_ref, 0u, (unsigned)live, (unsigned)(live >> 32));
    int sp = 1;
    while (sp > 0) {
        --sp;
        const uint4 e = s[sp];
        const int ref = __builtin_amdgcn_readfirstlane((int)e.x);
        const int lev = __builtin_amdgcn_readfirstlane((int)e.y);
        const uint64_t mask = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.w) << 32) |
                              (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.z);
        const bool is_root = !BATCH && ref == kRootRef;
        float4 q = nodes4[is_root ? 0 : ref];  // index kept in bounds for the root entry
        if (is_root) q = root;
        const float dx = q.x - me.x;
        const float dy = q.y - me.y;
        const float dz = q.z - me.z;
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        uint64_t takem, omask;
        if (ref < n) {
            takem = mask;
            omask = 0;
        } else {
            const float crit = ldexpf(crit0, -2 * lev);
            const uint64_t ok = __ballot(crit < d2);
            takem = mask & ok;
            omask = mask & ~ok;
        }
        if (COUNT) {
            if (__builtin_amdgcn_inverse_ballot_w64(takem)) cnt += (ref == p) ? 0 : 1;
        }
        if (!SOFT) takem &= __ballot(d2 > 0.f);
        if (__builtin_amdgcn_inverse_ballot_w64(takem)) {
            const float r2 = d2 + eps2;
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float f = q.w * ri * (ri * ri);
            px = fmaf(f, dx, px);
            py = fmaf(f, dy, py);
            pz = fmaf(f, dz, pz);
        }
        if (KAHAN) {
            if (++tile == 16) {
                kahan_add(sx, cx, px); kahan_add(sy, cy, py); kahan_add(sz, cz, pz);
                px = py = pz = 0.f;
                tile = 0;
            }
        }
        if (omask) {
            const int base = is_root ? 0 : 8 * ref;
            const int ch = lane < 8 ? tchild[base + lane] : -1;
            const uint64_t vm = __ballot(ch != -1);
            const int nch = __popcll(vm);
            if (sp + nch > kForceStack) {
                if (lane == 0) atomicOr(err, kErrStack);
                break;
            }
            if (ch != -1) {
                // slot 0 ends on top of the stack, so children are visited in slot order
                const int above = __popcll(vm & ~((2ull << lane) - 1ull));
                s[sp + above] = make_uint4((unsigned)ch, (unsigned)(lev + 1), (unsigned)omask,
                                           (unsigned)(omask >> 32));
            }
            sp += nch;
        }
    }
    if (KAHAN) {
        kahan_add(sx, cx, px); kahan_add(sy, cy, py); kahan_add(sz, cz, pz);
    } else {
        sx = px; sy = py; sz = pz;
    }
    if (valid) {
        const int i = perm[p];
        ax[i] = G * sx;
        ay[i] = G * sy;
        az[i] = G * sz;
        if (COUNT) counts[i] = cnt;
    }
}

struct ForceArgs {
    int n, ngroups;
    const float4 *nodes4;
    const int32_t *tchild, *perm;
    const DeviceState *st;
    float theta, eps2, G;
    float *ax, *ay, *az;
    int64_t *counts;
    int *err;
    const int4 *groups;
    const float *sys_theta;
    const double *sys_ext;
};

template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH, int W>
static void launch_one(const ForceArgs &a, hipStream_t s)
{
    const int nblk = (a.ngroups + W - 1) / W;
    force_kernel<KAHAN, COUNT, SOFT, BATCH, W><<<nblk, 64 * W, 0, s>>>(
        a.n, a.ngroups, a.nodes4, a.tchild, a.perm, a.st, a.theta, a.eps2, a.G, a.ax, a.ay, a.az, a.counts, a.err,
        a.groups, a.sys_theta, a.sys_ext);
}

template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH>
static void launch_w(const ForceArgs &a, hipStream_t s, int waves)
{
    if (waves <= 2) launch_one<KAHAN, COUNT, SOFT, BATCH, 2>(a, s);
    else if (waves >= 8) launch_one<KAHAN, COUNT, SOFT, BATCH, 8>(a, s);
    else launch_one<KAHAN, COUNT, SOFT, BATCH, 4>(a, s);
}

template <bool BATCH>
static void dispatch(const ForceArgs &a, hipStream_t s, bool kahan, int waves)
{
    const bool soft = a.eps2 > 0.f;
    if (a.counts) {
        if (soft) launch_w<false, true, true, BATCH>(a, s, waves);
        else launch_w<false, true, false, BATCH>(a, s, waves);
    } else if (kahan) {
        if (soft) launch_w<true, false, true, BATCH>(a, s, waves);
        else launch_w<true, false, false, BATCH>(a, s, waves);
    } else {
        if (soft) launch_w<false, false, true, BATCH>(a, s, waves);
        else launch_w<false, false, false, BATCH>(a, s, waves);
    }
}

static ForceArgs single_args(nbx_context *c, float *ax, float *ay, float *az, float g, int64_t *counts)
{
    ForceArgs a{};
    a.n = c->n;
    a.ngroups = (c->n + 63) / 64;
    a.nodes4 = c->nodes4; a.tchild = c->tchild; a.perm = c->perm; a.st = c->st;
    a.theta = c->p.theta;
    a.eps2 = c->p.softening * c->p.softening;
    a.G = g;
    a.ax = ax; a.ay = ay; a.az = az; a.counts = counts;
    a.err = &c->st->err;
    return a;
}

void launch_force(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int64_t *counts)
{
    if (c->n <= 0) return;
    dispatch<false>(single_args(c, ax, ay, az, g, counts), s, c->p.kahan != 0, c->force_waves);
}

void launch_force_waves(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int waves)
{
    if (c->n <= 0) return;
    dispatch<false>(single_args(c, ax, ay, az, g, nullptr), s, c->p.kahan != 0, waves);
}

void launch_batch_force(nbx_batch *b, hipStream_t s, int64_t *counts, int waves)
{
    ForceArgs a{};
    a.n = b->N;
    a.ngroups = b->ngroups;
    a.nodes4 = b->nodes4; a.tchild = b->tchild; a.perm = b->perm; a.st = b->st;
    a.theta = b->p.theta;
    a.eps2 = b->p.softening * b->p.softening;
    a.G = b->p.gravity;
    a.ax = counts ? b->acc_scratch : b->ax;
    a.ay = counts ? b->acc_scratch + b->N : b->ay;
    a.az = counts ? b->acc_scratch + 2 * (size_t)b->N : b->az;
    a.counts = counts;
    a.err = &b->st->err;
    a.groups = b->groups;
    a.sys_theta = b->theta;
    a.sys_ext = b->ext;
    dispatch<true>(a, s, b->p.kahan != 0, waves);
}

}  // namespace nbx
