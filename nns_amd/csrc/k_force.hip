// k_force.hip -- Barnes-Hut accelerations: 64-lane wavefront traversal.
//
// One wave owns a group of up to 64 consecutive bodies in key order, one body
// per lane, and walks the tree depth first.  The unit of the walk is a frame:
// an opened cell together with the mask of the lanes that opened it.  A frame
// names a block of the frame layout (the cell's non-empty children, compacted,
// with chains of one-child cells collapsed to their bottom cell), packed: the
// next block's records follow the last child.  Eight records from the block's
// first -- 8 float4 (x, y, z, m) and 8 int2 descriptors, 192 bytes -- come in
// through the scalar cache with three s_load_dwordx16 once per frame, and only
// the block's own children are stepped through, with their data already in
// SGPRs, so the only dependent memory access of a frame is that one load;
// frames live in an LDS stack of {tag = first record << 4 | children, 0, lane
// mask low, lane mask high}.
//
// Each lane applies its own opening test to a cell child (l/D < theta written
// as crit_L + eps^2 < |d|^2 + eps^2 with crit_L = (l_L/theta)^2, the reference
// CPU test of 3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:64-67; the left side
// comes with the child's descriptor, the right side is the interaction's own
// |d|^2 + eps^2, so the test costs one compare).  The test is a ballot, so lane
// masks stay in SGPRs: lanes of mask & accept add the cell's monopole (softened:
// computed on every lane and zeroed by a select on the others; unsoftened and
// counting variants: under an exec mask of those lanes), lanes of mask & ~accept
// push the child's frame.  Every lane therefore sees exactly its own
// per-particle interaction list; the wave pays for the union of the 64 lists.
// Bodies are always direct (Interaction.cpp:60-63); a zero distance contributes
// nothing (:28-30).
//
// Interaction: a += m d (|d|^2 + eps^2)^-3/2, d = q - x (the 20-flop form of
// GPU Gems 3 ch. 31), with the raw v_rsq_f32 (|d|^2 + eps^2 is never denormal
// when eps > 0; the unsoftened variant masks d = 0).  With Kahan on, monopoles
// go into a plain partial that is folded into a compensated total every 8
// frames.  Bookkeeping that is not per child (stack bound, Kahan) is per frame.
//
// Each XCD (dispatch is round-robin over the 8 XCDs) walks a range of
// consecutive groups, so that neighbouring walks share its L2; the ranges are
// sized by the child steps each group took in the previous launch, so that the
// XCDs finish together (a split by group count leaves the launch waiting for
// the XCD with the most work: 15 % over the mean for an ensemble of uneven
// systems, 6 % for a centrally condensed sphere).
// The launch's arguments that a group needs only when its walk starts or ends
// are read from a job record in device memory (ForceJob), so that they hold
// no scalar registers during the walk.
//
// Single system: group g is bodies [64g, 64g+64) of the group order (gorder:
// sorted positions ordered by coarse Hilbert key) and the walk starts from
// froot[0].  Ensemble: group g comes from a table {first, count, system};
// the walk starts at froot[system] (thresholds of the system's own theta and
// extent are in the descriptors).
//
// The counting variant (COUNT) also records, per group, the child steps (one
// step tests one child for the 64 lanes) and the frames of the walk.
#include "nbx_device.h"

namespace nbx {

// The walk of group g by the calling wave (all 64 lanes active), with the
// wave's LDS stack s.  The launch's other arguments come from its job record
// jb, read when the walk starts and again when it ends (the empty asm with a
// memory clobber keeps the compiler from holding them through the walk).
template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH>
static __device__ __forceinline__ void walk_group(int g, int lane, uint4 *s, const float4 *__restrict__ fnode,
                                                  const int2 *__restrict__ fmeta, float eps2, const ForceJob *jb)
{
    asm volatile("" ::: "memory");
    int first, cnt_bodies, sys;
    if (BATCH) {
        const int4 gr = jb->groups[g];
        first = gr.x;
        cnt_bodies = gr.y;
        sys = gr.z;
    } else {
        first = g * 64;
        cnt_bodies = min(64, jb->n - first);
        sys = 0;
    }
    const bool valid = lane < cnt_bodies;
    const int p = valid ? jb->gorder[first + lane] : 0;  // the lane's body, by sorted position
    const float4 me = valid ? jb->nodes4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t live = __ballot(valid);

    float sx = 0.f, sy = 0.f, sz = 0.f;  // compensated totals
    float cx = 0.f, cy = 0.f, cz = 0.f;  // compensations
    float px = 0.f, py = 0.f, pz = 0.f;  // partials of the current tile
    int64_t cnt = 0, nsteps = 0, nframes = 0;
    int work = 1;  // child steps of the walk (the root's and every frame's children)

    // One child step: every lane of `mask` meets child q with descriptor
    // {tag, crit}; returns the lanes that open it.  A body's crit is -1, so it
    // is always taken and never opened.
    auto step = [&](const float4 q, const int tag, const float crit, const uint64_t mask) -> uint64_t {
        const float dx = q.x - me.x;
        const float dy = q.y - me.y;
        const float dz = q.z - me.z;
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));  // |d|^2 + eps^2
        const uint64_t ok = __ballot(crit < d2);                         // crit = crit_L + eps^2
        uint64_t takem = mask & ok;
        const uint64_t omask = mask & ~ok;
        if (COUNT) {
            if (__builtin_amdgcn_inverse_ballot_w64(takem)) cnt += ((tag & 15) == 0 && (tag >> 4) == p) ? 0 : 1;
        }
        if (!SOFT) takem &= __ballot(d2 > 0.f);
        if (SOFT && !COUNT) {
            // |d|^2 + eps^2 > 0 on every lane, so the monopole is computed for all
            // lanes and zeroed by one select for those that do not take the child:
            // no exec-mask branch, fewer scalar instructions per step
            const float ri = __builtin_amdgcn_rsqf(d2);
            float f = q.w * ri * (ri * ri);
            f = __builtin_amdgcn_inverse_ballot_w64(takem) ? f : 0.f;
            px = fmaf(f, dx, px);
            py = fmaf(f, dy, py);
            pz = fmaf(f, dz, pz);
        } else if (__builtin_amdgcn_inverse_ballot_w64(takem)) {  // exec = takem: no select, skipped when empty
            asm volatile("");  // keeps the branch (no if-conversion into selects of the partials)
            const float ri = __builtin_amdgcn_rsqf(d2);
            const float f = q.w * ri * (ri * ri);
            px = fmaf(f, dx, px);
            py = fmaf(f, dy, py);
            pz = fmaf(f, dz, pz);
        }
        return omask;
    };

    int sp = 0;
    {
        const int2 rd = jb->frootm[sys];
        const int rt = __builtin_amdgcn_readfirstlane(rd.x);
        const float rc = __int_as_float(__builtin_amdgcn_readfirstlane(rd.y));
        const uint64_t om = step(jb->froot4[sys], rt, rc, live);
        if (COUNT) nsteps += 1;
        if (om) {
            if (lane == 0) s[0] = make_uint4((unsigned)rt, 0u, (unsigned)om, (unsigned)(om >> 32));
            sp = 1;
        }
    }
    // Frames run in tiles of kTile; the partials are folded after each tile.
    // The frame loop has one exit, so the accumulators keep their registers
    // from frame to frame (no copies between iterations).
    constexpr int kTile = KAHAN ? 8 : 0x7fffffff;
    bool overflow = false;
    do {
      for (int t = 0; t < kTile; ++t) {
        if (sp <= 0) break;
        // a frame pops one entry and pushes at most 8
        if (sp > kForceStack - 7) {
            overflow = true;
            sp = 0;
            break;
        }
        sp = __builtin_amdgcn_readfirstlane(sp - 1);
        const uint4 e = s[sp];
        const int tag = __builtin_amdgcn_readfirstlane((int)e.x);
        const int nch = tag & 15;
        const int off = tag >> 4;  // the block's first record
        const uint64_t mask = ((uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.w) << 32) |
                              (uint64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)e.z);
        work += nch;
        if (COUNT) {
            nsteps += nch;
            nframes += 1;
        }
        const float4 *bq = fnode + off;
        const int2 *bd = fmeta + off;
        float4 q[8];
        int2 d[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            q[j] = bq[j];
            d[j] = bd[j];
        }
        // Every load of the block is issued before the first use: one wait per frame.
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
        if (KAHAN) {  // fold the partials of the tile (about 40 children)
            kahan_add(sx, cx, px); kahan_add(sy, cy, py); kahan_add(sz, cz, pz);
            px = py = pz = 0.f;
        }
    } while (sp > 0);
    asm volatile("" ::: "memory");
    if (overflow && lane == 0) atomicOr(jb->err, kErrStack);
    if (!KAHAN) {
        sx = px; sy = py; sz = pz;
    }
    if (valid) {
        const int i = jb->perm[p];
        const float G = jb->G;
        jb->ax[i] = G * sx;
        jb->ay[i] = G * sy;
        jb->az[i] = G * sz;
        if (COUNT) jb->counts[i] = cnt;
    }
    if (lane == 0) jb->gcost[g] = work;  // the next launch splits the groups by it
    if (COUNT && jb->walk && lane == 0) {
        jb->walk[2 * (int64_t)g] = nsteps;
        jb->walk[2 * (int64_t)g + 1] = nframes;
    }
}

// Groups per XCD at most: 1.25 times the mean, so that the launch has a fixed
// size while the split follows the work.
static inline int xcd_cap(int ngroups) { return std::max((ngroups + 7) / 8, (5 * ngroups + 31) / 32); }

// Block b runs on XCD x = b mod 8 (dispatch is round-robin over the 8 XCDs) as
// that XCD's block i = b / 8: its waves walk groups fcut[x] + W i + wave, so an
// XCD walks consecutive groups (neighbouring walks share its L2) and its share
// of the groups is the one the setup kernel chose; blocks past the share exit.
template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH, int W>
__global__ void __launch_bounds__(64 * W) force_kernel(const float4 *__restrict__ fnode,
                                                       const int2 *__restrict__ fmeta, float eps2, DeviceState *st)
{
    __shared__ uint4 stk[W][kForceStack];
    if (latched(st)) return;
    const int x = blockIdx.x & 7;
    const int g = __builtin_amdgcn_readfirstlane(st->fcut[x] + (int)(blockIdx.x >> 3) * W + (int)(threadIdx.x >> 6));
    if (g < st->fcut[x + 1])
        walk_group<KAHAN, COUNT, SOFT, BATCH>(g, threadIdx.x & 63, stk[threadIdx.x >> 6], fnode, fmeta, eps2, &st->fj);
}

// A group's weight in the split: its recorded child steps, within [1, 2^24].
static __device__ inline int64_t group_cost(int32_t c) { return min(max(c, 1), 1 << 24); }

// Before each force launch: stores the launch's job record and splits its
// groups over the 8 XCDs (a kernel, not a copy, so that the force stage can be
// captured in a graph).  The split gives each XCD an equal share of the child
// steps that the previous launch over the same groups counted (gcost), within
// xcd_cap groups per XCD; without such a launch the shares are equal in groups.
// Thread t sums the costs of groups [t C, t C + C), C a multiple of 4 (read as
// int4), and the cuts fall on those chunk boundaries.  Every group is walked by
// exactly one wave whatever the split, so the results do not depend on it.
__global__ void __launch_bounds__(1024) force_setup_kernel(DeviceState *st, ForceJob jb, int cap)
{
    __shared__ int64_t wsum[16];
    __shared__ int ideal[9];
    const int G = jb.ngroups, t = threadIdx.x;
    const bool hist = st->fcost_n == G;
    const int C = ((G + 1023) / 1024 + 3) & ~3;
    const int a = min(G, t * C), b = min(G, a + C);
    int64_t sum = 0;
    if (hist) {  // batches of 8 int4 loads, all issued before the first sum
        int k = a;
        for (; k + 32 <= b; k += 32) {
            int4 v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = *reinterpret_cast<const int4 *>(jb.gcost + k + 4 * i);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                sum += group_cost(v[i].x) + group_cost(v[i].y) + group_cost(v[i].z) + group_cost(v[i].w);
        }
        for (; k + 4 <= b; k += 4) {
            const int4 v = *reinterpret_cast<const int4 *>(jb.gcost + k);
            sum += group_cost(v.x) + group_cost(v.y) + group_cost(v.z) + group_cost(v.w);
        }
        for (; k < b; ++k) sum += group_cost(jb.gcost[k]);
    }
    // inclusive scan of the per-thread sums: within each wave by shuffles, then
    // over the 16 wave totals
    const int lane = t & 63, wave = t >> 6;
    int64_t inc = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    if (t < 9) ideal[t] = (int)((int64_t)G * t / 8);
    __syncthreads();
    if (t == 0)
        for (int w = 1; w < 16; ++w) wsum[w] += wsum[w - 1];
    __syncthreads();
    inc += wave > 0 ? wsum[wave - 1] : 0;
    const int64_t total = wsum[15], lo = inc - sum;
    if (hist && total > 0)
        for (int x = 1; x < 8; ++x) {  // the chunk in which the prefix reaches x/8 of the steps
            const int64_t target = total * x / 8;
            if (target >= lo && target < inc) ideal[x] = target - lo < inc - target ? a : b;
        }
    __syncthreads();
    if (t == 0) {
        int cut = 0;
        st->fcut[0] = 0;
        for (int x = 0; x < 8; ++x) {  // monotone, at most cap per XCD, all G groups covered
            const int lo_b = max(cut, G - (7 - x) * cap), hi_b = min(G, cut + cap);
            cut = min(max(ideal[x + 1], lo_b), hi_b);
            st->fcut[x + 1] = cut;
        }
        st->fcost_n = G;
        st->fj = jb;
    }
}

struct ForceArgs {
    ForceJob jb;
    const float4 *fnode;
    const int2 *fmeta;
    DeviceState *st;
    float eps2;
};

template <bool KAHAN, bool COUNT, bool SOFT, bool BATCH, int W>
static void launch_one(const ForceArgs &a, hipStream_t s)
{
    const int cap = xcd_cap(a.jb.ngroups);
    force_setup_kernel<<<1, 1024, 0, s>>>(a.st, a.jb, cap);
    force_kernel<KAHAN, COUNT, SOFT, BATCH, W><<<8 * ((cap + W - 1) / W), 64 * W, 0, s>>>(a.fnode, a.fmeta, a.eps2,
                                                                                      a.st);
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
    if (a.jb.counts) {
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
    ForceJob &j = a.jb;
    j.n = c->n;
    j.ngroups = (c->n + 63) / 64;
    j.nodes4 = c->nodes4; j.gorder = c->gorder; j.froot4 = c->froot4; j.frootm = c->frootm;
    j.perm = c->perm;
    j.G = g;
    j.ax = ax; j.ay = ay; j.az = az; j.counts = counts;
    j.err = &c->st->err;
    j.gcost = c->gcost;
    a.fnode = c->fnode; a.fmeta = c->fmeta; a.st = c->st;
    a.eps2 = layout_eps2(c->p);
    return a;
}

void launch_force(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int64_t *counts,
                  int64_t *walk)
{
    if (c->n <= 0) return;
    if (frame_layout_stale(c)) launch_frame_layout(c, s);  // thresholds of the current theta and eps
    ForceArgs a = single_args(c, ax, ay, az, g, counts);
    a.jb.walk = walk;
    dispatch<false>(a, s, c->p.kahan != 0, c->force_waves);
}

void launch_force_waves(nbx_context *c, hipStream_t s, float *ax, float *ay, float *az, float g, int waves)
{
    if (c->n <= 0) return;
    if (frame_layout_stale(c)) launch_frame_layout(c, s);
    dispatch<false>(single_args(c, ax, ay, az, g, nullptr), s, c->p.kahan != 0, waves);
}

void launch_batch_force(nbx_batch *b, hipStream_t s, int64_t *counts, int waves, int64_t *walk)
{
    ForceArgs a{};
    ForceJob &j = a.jb;
    j.walk = walk;
    j.n = b->N;
    j.ngroups = b->ngroups;
    j.nodes4 = b->nodes4; j.gorder = b->gorder; j.froot4 = b->froot4; j.frootm = b->frootm;
    j.perm = b->perm;
    j.G = b->p.gravity;
    j.ax = counts ? b->acc_scratch : b->ax;
    j.ay = counts ? b->acc_scratch + b->N : b->ay;
    j.az = counts ? b->acc_scratch + 2 * (size_t)b->N : b->az;
    j.counts = counts;
    j.err = &b->st->err;
    j.groups = b->groups;
    j.gcost = b->gcost;
    a.fnode = b->fnode; a.fmeta = b->fmeta; a.st = b->st;
    a.eps2 = layout_eps2(b->p);
    dispatch<true>(a, s, b->p.kahan != 0, waves);
}

}  // namespace nbx
