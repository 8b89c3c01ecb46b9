// k_force.hip -- Barnes-Hut accelerations: 64-lane wavefront traversal.
//
// One wave owns a group of 64 consecutive bodies in key order, one body per
// lane.  The wave walks the tree depth first with an LDS stack of entries
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
#include "nbx_internal.h"

namespace nbx {

constexpr int kForceWaves = 4;

__device__ inline int xcd_swizzle(int b, int nblk)
{
    const int q = nblk / 8, r = nblk % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

__device__ inline void kahan_add(float &sum, float &comp, float v)
{
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

template <bool KAHAN, bool COUNT, bool SOFT>
__global__ void __launch_bounds__(64 * kForceWaves) force_kernel(
    int n, const float4 *__restrict__ nodes4, const int32_t *__restrict__ tchild,
    const int32_t *__restrict__ perm, const DeviceState *st, float theta, float eps2, float G,
    float *__restrict__ ax, float *__restrict__ ay, float *__restrict__ az, int64_t *__restrict__ counts,
    int *__restrict__ err)
{
    __shared__ uint4 stk[kForceWaves][kForceStack];
    if (__hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = xcd_swizzle(blockIdx.x, gridDim.x) * kForceWaves + wave;
    const int ngroups = (n + 63) >> 6;
    if (g >= ngroups) return;

    // crit_0 = (l_0/theta)^2, l_0 = side of the root cell of the key grid
    const double lc0 = ldexp(st->ext[3] / kMaskCoord, kKeyLevels);
    const double r0 = lc0 / (double)theta;
    const float crit0 = (float)(r0 * r0);
    const float4 root = st->root4;

    const int p = g * 64 + lane;
    const bool valid = p < n;
    const float4 me = valid ? nodes4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t live = __ballot(valid);

    float sx = 0.f, sy = 0.f, sz = 0.f;  // compensated totals
    float cx = 0.f, cy = 0.f, cz = 0.f;  // compensations
    float px = 0.f, py = 0.f, pz = 0.f;  // partials of the current tile
    int64_t cnt = 0;
    int tile = 0;

    uint4 *s = stk[wave];
    if (lane == 0) s[0] = make_uint4((unsigned)kRootRef, 0u, (unsigned)live, (unsigned)(live >> 32));
    int sp = 1;
    while (sp > 0) {
        --sp;
        const uint4 e = s[sp];
        const int ref = __builtin_amdgcn_readfirstlane((int)e.x);
        const int lev = __builtin_amdgcn_readfirstlane((int)e.y);
        const uint64_t mask = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.w) << 32) |
                              (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.z);
        const bool is_root = ref == kRootRef;
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

template <bool KAHAN, bool COUNT, bool SOFT>
static void launch_variant(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g,
                           int64_t *counts, float eps2)
{
    const int n = c->n;
    const int ngroups = (n + 63) / 64;
    const int nblk = (ngroups + kForceWaves - 1) / kForceWaves;
    force_kernel<KAHAN, COUNT, SOFT><<<nblk, 64 * kForceWaves, 0, s>>>(
        n, c->nodes4, c->tchild, c->perm, c->st, c->p.theta, eps2, g, ax, ay, az, counts, &c->st->err);
}

void launch_force(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int64_t *counts)
{
    if (c->n <= 0) return;
    const float eps2 = c->p.softening * c->p.softening;
    const bool soft = eps2 > 0.f;
    if (counts) {
        if (soft) launch_variant<false, true, true>(c, s, ax, ay, az, g, counts, eps2);
        else launch_variant<false, true, false>(c, s, ax, ay, az, g, counts, eps2);
    } else if (c->p.kahan) {
        if (soft) launch_variant<true, false, true>(c, s, ax, ay, az, g, nullptr, eps2);
        else launch_variant<true, false, false>(c, s, ax, ay, az, g, nullptr, eps2);
    } else {
        if (soft) launch_variant<false, false, true>(c, s, ax, ay, az, g, nullptr, eps2);
        else launch_variant<false, false, false>(c, s, ax, ay, az, g, nullptr, eps2);
    }
}

}  // namespace nbx
