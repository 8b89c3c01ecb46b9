// k_tree.hip -- tree construction, bottom-up summary and the traversal layout.
//
// Cells are prefixes of the 63-bit Morton key.  A cell at level L selects the
// slot of a body by key digit L (bits 62-3L .. 60-3L), so slot order is key
// order and the bodies of any cell are a contiguous range of the sorted keys.
// Single system: the root is node 0 with its slots at child[0..7] (the
// reference's convention, src/Kernels.cu:195,225) and cells take ids n, n+1, ...
// Ensemble: system s has the root cell N + s and cells follow from N + S.
//
// Construction is split along the MI355X memory hierarchy.  Each XCD has its
// own L2, so a fence or atomic at device scope costs a trip past L2; those are
// kept to a handful per subtree:
//   1. top cells -- cells holding more than kSubtreeMax bodies -- are found
//      from the sorted keys by binary search (one thread per sample of
//      kSubtreeMax bodies) and wired with their child ranges;
//   2. every child range of a top cell with 2..kSubtreeMax bodies is a subtree
//      built by ONE wave from the shared-digit counts of neighbouring keys:
//      cell ids from a prefix sum, parents from previous-smaller-value queries,
//      one plain store per child slot (subtree_build_kernel).  Every subtree's
//      exact cell count is known first (from the shared-digit counts), and one
//      scan of the counts gives each its block of cell ids.
//      The lock-based insertion (compare-and-swap on child slots, -2 = locked,
//      at workgroup scope; subtree_insert_kernel) builds the same cells when
//      nbx_params.build_locked is set;
//   3. the summary runs level by level: subtrees inside their workgroup, then
//      the top cells.  Every cell is summed by one thread over its
//      children in slot order, in FP64: results do not depend on scheduling.
#include "nbx_device.h"

#include <algorithm>

namespace nbx {

// Segments: one system ([0, n), root node 0) or an ensemble (off[], roots n + s).
struct Segs {
    const int32_t *off;       // nullptr: one segment [0, n)
    const int32_t *samp_off;  // nullptr: its samples start at 0
    int nseg, n;
    int batch;                // roots at n + s instead of node 0
};

__device__ inline int seg_lo(const Segs &g, int s) { return g.off ? g.off[s] : 0; }
__device__ inline int seg_hi(const Segs &g, int s) { return g.off ? g.off[s + 1] : g.n; }
__device__ inline int samp_lo(const Segs &g, int s) { return g.samp_off ? g.samp_off[s] : 0; }
__device__ inline int seg_root(const Segs &g, int s) { return g.batch ? g.n + s : 0; }
__device__ inline int seg_of_sample(const Segs &g, int k)
{
    if (!g.samp_off) return 0;
    int lo = 0, hi = g.nseg;  // last s with samp_off[s] <= k
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (g.samp_off[mid] <= k) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ inline uint64_t prefix_of(uint64_t key, int L) { return key >> (63 - 3 * L); }

// First index in [lo, hi) whose level-L prefix is >= P (keys sorted).
__device__ inline int lower_prefix(const uint64_t *k, int lo, int hi, uint64_t P, int L)
{
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (prefix_of(k[mid], L) < P) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ inline int upper_prefix(const uint64_t *k, int lo, int hi, uint64_t P, int L)
{
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (prefix_of(k[mid], L) <= P) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// First index in [lo, hi) whose digit L is >= d (all keys share the first L digits).
__device__ inline int lower_digit(const uint64_t *k, int lo, int hi, int d, int L)
{
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (digit_at(k[mid], L) < d) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// lcp[k] = number of leading key digits shared by sorted keys k and k+1 (21 if equal).
__global__ void __launch_bounds__(256) lcp_kernel(int n, const uint64_t *__restrict__ skeys,
                                                  uint8_t *__restrict__ lcp, DeviceState *st)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->nsub = 0;
        st->ntop = 0;
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n - 1; k += blockDim.x * gridDim.x) {
        const uint64_t x = (skeys[k] ^ skeys[k + 1]) & ~(1ull << 63);
        lcp[k] = x == 0 ? (uint8_t)kKeyLevels : (uint8_t)((62 - (63 - __clzll((long long)x))) / 3);
    }
}

// Top cells: one lane per (sample q, level L); q is every kSubtreeMax-th sorted
// position of a segment.  The level-L cell through q is [a, b), found by two
// binary searches over the segment; it is a top cell when it holds more than
// kSubtreeMax bodies, and q owns it when no earlier sample lies in it.  The
// levels of a sample are searched side by side, not one after another.  The
// table is level-major (entry L * nsamples + k), so the scan of its flags
// numbers the top cells level by level: the cells of level L are the ids
// base0 + [id[L * nsamples], id[(L + 1) * nsamples]).
__global__ void __launch_bounds__(256) top_cells_kernel(Segs g, int nsamples, const uint64_t *__restrict__ skeys,
                                                        uint32_t *__restrict__ flag, uint32_t *__restrict__ id,
                                                        int32_t *__restrict__ ta, int32_t *__restrict__ tb,
                                                        const DeviceState *st)
{
    if (latched(st)) return;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = gt >> 5, L = gt & 31;
    if (k >= nsamples || L >= kLevels) return;
    const int e = L * nsamples + k;
    flag[e] = 0u;
    id[e] = 0u;
    if (L == 0) return;  // the root is wired separately
    const int s = seg_of_sample(g, k);
    const int lo = seg_lo(g, s), hi = seg_hi(g, s);
    const int q = lo + (k - samp_lo(g, s)) * kSubtreeMax;
    const uint64_t P = prefix_of(skeys[q], L);
    const int a = lower_prefix(skeys, lo, q + 1, P, L);
    const int b = upper_prefix(skeys, q, hi, P, L);
    if (b - a > kSubtreeMax && (q - kSubtreeMax < a || q == lo)) {
        flag[e] = 1u;
        id[e] = 1u;
        ta[e] = a;
        tb[e] = b;
    }
}

// Wires every top cell (and every root): child ranges by digit L; a range of
// one body is a leaf, of more than kSubtreeMax a top cell, otherwise a subtree
// queued for the subtree build.  Eight lanes per entry, lane d for digit d: each
// finds the end of its digit's range by one binary search over the cell's
// range and takes its start from the lane before (the searches run side by
// side, not one after another).  Thread 0 also sets *index past the top cells.
__global__ void __launch_bounds__(256) top_wire_kernel(
    Segs g, int nsamples, const uint64_t *__restrict__ skeys, const int32_t *__restrict__ perm, int *child,
    int32_t *__restrict__ cell_level, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ id,
    const int32_t *__restrict__ ta, const int32_t *__restrict__ tb, int4 *__restrict__ sub, int sub_max,
    int base0, int m, int *index, int *__restrict__ start, DeviceState *st)
{
    if (latched(st)) return;
    const int total = nsamples * kLevels;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = gt >> 3;  // entry: one (level, sample) of the top table, or one root
    const int d = gt & 7;   // digit, slot
    if (gt == 0) {
        const int ntop = (int)(id[total - 1] + flag[total - 1]);
        st->ntop = ntop;
        *index = base0 + ntop;
        if ((int64_t)base0 + ntop > m) atomicOr(&st->err, kErrCapacity);
    }
    int c, L, a, b, s;
    if (t < total) {
        if (!flag[t]) return;
        L = t / nsamples;
        const int k = t - L * nsamples;
        c = base0 + (int)id[t];
        a = ta[t];
        b = tb[t];
        s = seg_of_sample(g, k);
        if (c >= m) return;  // capacity error latched by thread 0
        if (d == 0) {
            cell_level[c] = L;
            start[c] = a;
        }
        if (L == kKeyLevels) {  // more than kSubtreeMax bodies on one key
            if (d == 0) atomicOr(&st->err, kErrBucket);
            child[8 * (int64_t)c + d] = -1;
            return;
        }
    } else if (t < total + g.nseg) {
        s = t - total;
        c = seg_root(g, s);
        L = 0;
        a = seg_lo(g, s);
        b = seg_hi(g, s);
        if (g.batch && d == 0) start[c] = a;  // the single system's root is node 0, not a cell
    } else {
        return;
    }
    // the 8 lanes of an entry are consecutive lanes of one wave and all get here
    const int lo = seg_lo(g, s);
    const int db = lower_digit(skeys, a, b, d + 1, L);
    const int up = __shfl_up(db, 1, 8);
    const int da = d == 0 ? a : up;
    const int cnt = db - da;
    const int64_t slot = 8 * (int64_t)c + d;
    if (cnt == 0) {
        child[slot] = -1;
    } else if (cnt == 1) {
        child[slot] = perm[da];
    } else if (cnt > kSubtreeMax) {
        const int kk = samp_lo(g, s) + (da - lo + kSubtreeMax - 1) / kSubtreeMax;
        child[slot] = base0 + (int)id[(L + 1) * nsamples + kk];
    } else {
        child[slot] = -1;  // set by the subtree build
        const int i = atomicAdd(&st->nsub, 1);
        if (i < sub_max) sub[i] = make_int4(da, db, L + 1, (int)slot);
        else atomicOr(&st->err, kErrCapacity);
    }
}

// Cell ids of the subtrees.  A subtree [a, b) of sorted bodies with its root at
// level L0 has one cell per run of two or more bodies sharing L digits, L0 <= L
// <= 21; body a + i starts the runs of levels P_i < L <= l_i (l_i = lcp[a + i],
// P_i = l_{i-1}, P_0 = L0 - 1; the last body starts none), so the subtree has
//     (lcp[a] - L0 + 1) + sum over a < k < b - 1 of max(0, lcp[k] - lcp[k - 1])
// cells.  subtree_cells_kernel counts them, one wave per subtree (a job the
// builds reject counts 0); subtree_base_kernel turns the counts into the
// exclusive sum cbase[] (cbase[nsub] the total) and sets *index past the last
// subtree cell, so subtree j owns the ids from base0 + ntop + cbase[j] (after
// the top cells).  Each build workgroup reads its base instead of taking it from
// one shared counter: the device-scope atomics of a shared counter are
// serialised, one per subtree.  A build that counts other than cbase[j + 1] -
// cbase[j] cells for its subtree latches kErrCapacity and leaves it unbuilt, so
// no two subtrees can share an id.
__global__ void __launch_bounds__(256) subtree_cells_kernel(const int4 *__restrict__ sub, int sub_max,
                                                            const uint8_t *__restrict__ lcp,
                                                            uint32_t *__restrict__ cbase, const DeviceState *st)
{
    if (latched(st)) return;
    const int nsub = min(__hip_atomic_load(&st->nsub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), sub_max);
    const int lane = threadIdx.x & 63;
    const int waves = (int)(gridDim.x * blockDim.x) >> 6;
    for (int job = (int)(blockIdx.x * blockDim.x + threadIdx.x) >> 6; job < nsub; job += waves) {
        const int4 jb = sub[job];
        const int a = jb.x, cnt = jb.y - jb.x, L0 = jb.z;
        int c = 0;
        if (cnt >= 2 && cnt <= kSubtreeMax && L0 >= 1 && L0 <= kKeyLevels)
            for (int k = a + lane; k < a + cnt - 1; k += 64)
                c += k == a ? (int)lcp[a] - L0 + 1 : max(0, (int)lcp[k] - (int)lcp[k - 1]);
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) cbase[job] = (uint32_t)c;
    }
}

// One workgroup of 1024 threads: 16 counts per thread per pass.
__global__ void __launch_bounds__(1024) subtree_base_kernel(uint32_t *__restrict__ cbase, int sub_max, int base0,
                                                            int m, int *index, DeviceState *st)
{
    __shared__ uint32_t sh[16];
    if (latched(st)) return;
    const int nsub = min(__hip_atomic_load(&st->nsub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), sub_max);
    constexpr int V = 16;
    uint32_t carry = 0;
    for (int t0 = 0; t0 < nsub; t0 += V * (int)blockDim.x) {
        const int i0 = t0 + V * (int)threadIdx.x;
        uint32_t v[V], s = 0;
#pragma unroll
        for (int u = 0; u < V; ++u) {
            v[u] = i0 + u < nsub ? cbase[i0 + u] : 0u;
            s += v[u];
        }
        uint32_t tot;
        uint32_t run = carry + block_exclusive_scan(s, sh, &tot);
#pragma unroll
        for (int u = 0; u < V; ++u) {
            if (i0 + u < nsub) cbase[i0 + u] = run;
            run += v[u];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        cbase[nsub] = carry;  // so that subtree j has cbase[j + 1] - cbase[j] cells for every j
        const int64_t end = (int64_t)base0 + st->ntop + carry;
        if (end > m) atomicOr(&st->err, kErrCapacity);
        else *index = (int)end;
    }
}

__device__ inline int ld_slot_wg(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void st_slot_wg(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline int cas_slot_wg(int *p, int expect, int desired)
{
    __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expect;
}

__device__ inline int block_max(int v, int *red)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    int s = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = max(s, red[w]);
    __syncthreads();
    return s;
}

__device__ inline int block_sum(int v, int *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    int s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
}

// Subtree build by insertion (nbx_params.build_locked = 1): persistent
// workgroups take subtrees from the queue.  Inside a
// subtree, the bodies are inserted by lock-based descent along their key digits
// (one body per thread per round, bodies strided over the range so that the
// lanes of a wave start in different branches):
//   cell (>= n)  descend;   -1  CAS -1 -> body;   -2  locked, retry next iteration;
//   body         CAS body -> -2 (lock); the winner creates the chain of cells down
//                to the first level where the two keys differ (a level-21 bucket if
//                they are equal), then publishes the first cell with a release store
//                (unlock).
// No lane waits inside an iteration on another lane, so a wave whose lanes
// contend for one slot always makes progress.  All atomics and fences are at
// workgroup scope: only this workgroup touches the subtree.
__global__ void __launch_bounds__(256) subtree_insert_kernel(
    int n, int m, const uint64_t *__restrict__ skeys, const int32_t *__restrict__ perm,
    const uint64_t *__restrict__ keys, const uint8_t *__restrict__ lcp, int *child, int base0,
    const uint32_t *__restrict__ cbase, int32_t *__restrict__ cell_level, int4 *__restrict__ sub, int sub_max,
    int4 *__restrict__ subcells, DeviceState *st)
{
    __shared__ int s_job, s_base, s_next, s_bad;
    __shared__ int red[4];
    const int tid = threadIdx.x;
    const int nsub = min(__hip_atomic_load(&st->nsub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), sub_max);
    const int top = base0 + __hip_atomic_load(&st->ntop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int next = blockIdx.x;  // workgroups stride over the jobs, like the summary's
    for (;;) {
        // Thread 0 alone reads the error latch (which other workgroups of this
        // launch may raise at any time) and names the job; every wave reads its
        // decision from LDS after the barrier, so the workgroup continues or
        // leaves as a whole.  A wave that left on its own read would let the
        // others pass the barrier and read an s_job nobody wrote.
        if (tid == 0) s_job = latched(st) ? nsub : next;
        next += gridDim.x;
        __syncthreads();
        const int job = s_job;
        __syncthreads();
        if (job >= nsub) break;
        const int4 jb = sub[job];
        const int a = jb.x, b = jb.y, L0 = jb.z, slot = jb.w;
        // exact cell count of the subtree from the common-prefix lengths
        int contrib = 0, deep = L0;
        for (int k = a + tid; k < b - 1; k += blockDim.x) {
            contrib += (k == a) ? (int)lcp[a] - L0 + 1 : max(0, (int)lcp[k] - (int)lcp[k - 1]);
            deep = max(deep, (int)lcp[k]);
        }
        const int ks = block_sum(contrib, red);
        const int Lmax = block_max(deep, red);
        if (tid == 0) {
            const int base = top + (int)cbase[job];  // ids reserved by subtree_base_kernel
            s_bad = (int64_t)base + ks > (int64_t)m || ks != (int)(cbase[job + 1] - cbase[job]);
            if (s_bad) atomicOr(&st->err, kErrCapacity);
            s_base = base;
            s_next = 1;
            if (!s_bad) {
                for (int d = 0; d < 8; ++d) child[8 * (int64_t)base + d] = -1;
                cell_level[base] = L0;
                child[slot] = base;
                subcells[job] = make_int4(base, ks, L0, Lmax);
            } else {
                subcells[job] = make_int4(0, 0, L0, L0 - 1);
            }
        }
        __syncthreads();
        if (s_bad) continue;
        const int base = s_base, root = base;
        const int count = b - a;
        const int rounds = (count + blockDim.x - 1) / blockDim.x;
        for (int r = 0; r < rounds; ++r) {
            const int off = r + tid * rounds;
            bool done = off >= count;
            const int p = a + off;
            const int body = done ? 0 : perm[p];
            const uint64_t key = done ? 0ull : skeys[p];
            int node = root, L = L0;
            while (__ballot(!done)) {
                if (done) continue;
                if (L == kKeyLevels) {
                    bool placed = false, full = true;
                    for (int s = 0; s < 8 && !placed; ++s) {
                        int ch = ld_slot_wg(&child[8 * (int64_t)node + s]);
                        if (ch == -1) {
                            full = false;
                            if (cas_slot_wg(&child[8 * (int64_t)node + s], -1, body) == -1) placed = true;
                        }
                    }
                    if (placed) {
                        done = true;
                    } else if (full) {
                        atomicOr(&st->err, kErrBucket);
                        done = true;
                    }
                    continue;
                }
                const int64_t addr = 8 * (int64_t)node + digit_at(key, L);
                const int ch = ld_slot_wg(&child[addr]);
                if (ch >= n) {
                    node = ch;
                    ++L;
                } else if (ch == -1) {
                    if (cas_slot_wg(&child[addr], -1, body) == -1) done = true;
                } else if (ch >= 0) {
                    if (cas_slot_wg(&child[addr], ch, -2) == ch) {
                        const uint64_t key2 = keys[ch];
                        int ld = L + 1;
                        while (ld < kKeyLevels && digit_at(key, ld) == digit_at(key2, ld)) ++ld;
                        const int need = ld - L;  // cells at levels L+1 .. ld
                        const int old = atomicAdd(&s_next, need);
                        if (old + need > ks) {
                            atomicOr(&st->err, kErrCapacity);
                            st_slot_wg(&child[addr], ch);
                            done = true;
                            continue;
                        }
                        const int first = base + old;
                        const int4 neg = make_int4(-1, -1, -1, -1);
                        for (int j = 0; j < need; ++j) {
                            const int c = first + j;
                            const int lvl = L + 1 + j;
                            int4 *c4 = reinterpret_cast<int4 *>(child + 8 * (int64_t)c);
                            c4[0] = neg;
                            c4[1] = neg;
                            cell_level[c] = lvl;
                            if (lvl < ld) {
                                child[8 * (int64_t)c + digit_at(key, lvl)] = c + 1;
                            } else if (ld < kKeyLevels) {
                                child[8 * (int64_t)c + digit_at(key2, ld)] = ch;
                                child[8 * (int64_t)c + digit_at(key, ld)] = body;
                            } else {
                                child[8 * (int64_t)c + 0] = ch;  // bucket; index order restored in the summary
                                child[8 * (int64_t)c + 1] = body;
                            }
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                        st_slot_wg(&child[addr], first);  // publish + unlock
                        done = true;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Subtree build, without locks or atomics on child slots.  One wave builds one
// subtree [a, b) of sorted bodies whose root is at level L0 (a child range of
// a top cell, 2..kSubtreeMax bodies).  With l[i] the number of key digits that
// local bodies i and i+1 share (l[cnt-1] = L0 - 1 at the subtree's edge) and
// P_i = l[i-1] (P_0 = L0 - 1), the cells of the subtree are the runs of two or
// more bodies that share L digits, L = L0..21, and body i starts exactly the
// cells of levels P_i < L <= l[i].  Numbering cells by first body, then level,
//     id(i, L) = base + cs[i] + (L - P_i - 1),   cs = exclusive sum of max(0, l[i] - P_i),
// gives every cell its id from one prefix sum (base from one atomicAdd).
// The parent of cell (i, L), L > L0, is (i, L-1) when L-1 > P_i; otherwise the
// level-P_i cell that holds bodies i-1 and i, which starts at
//     s_i = 1 + the last j < i-1 with l[j] < P_i   (0 if none),
// a previous-smaller-value query answered by binary lifting over a table of
// range minima of l in LDS.  The deepest cell of body i is (i, l[i]) when
// l[i] > P_i, else the same (s_i, P_i).  A child takes the slot of its key
// digit at the parent's level; in a level-21 bucket (equal keys) the bodies
// take slots 0, 1, ... in sorted order, which is ascending index order.
// Each slot thus has exactly one writer: the wave sets the slots of its cells
// to -1, waits for those stores, and every cell and body stores itself into
// its parent.  Every id and slot is range-checked before its store.
constexpr int kLiftLevels = 10;  // 2^kLiftLevels >= kSubtreeMax
// Workgroups stride over the jobs.  With about one job each at N = 4M (16,400
// subtrees), the dispatcher starts a new workgroup wherever one has finished,
// which balances subtrees of 2 to 1024 bodies better than a few jobs per
// resident workgroup (0.141 ms with 4096 workgroups, 0.103 ms with 16384).
constexpr int kBuildBlocks = 16384;

__device__ inline int wave_exclusive_sum(int v, int *total)
{
    const int lane = threadIdx.x & 63;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

__global__ void __launch_bounds__(64) subtree_build_kernel(
    int m, const uint64_t *__restrict__ skeys, const int32_t *__restrict__ perm, const uint8_t *__restrict__ lcp,
    int *__restrict__ child, int base0, const uint32_t *__restrict__ cbase, int32_t *__restrict__ cell_level,
    int *__restrict__ start, const int4 *__restrict__ sub, int sub_max, int4 *__restrict__ subcells,
    DeviceState *st)
{
    constexpr int C = kSubtreeMax / 64;                   // bodies per lane at most
    __shared__ uint8_t s_min[kLiftLevels][kSubtreeMax];  // s_min[j][x] = min of l over [x, x + 2^j)
    __shared__ uint16_t s_cs[kSubtreeMax];                 // cs[i] (ids relative to base)
    if (latched(st)) return;
    const int lane = threadIdx.x;
    const int nsub = min(__hip_atomic_load(&st->nsub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), sub_max);
    const int top = base0 + __hip_atomic_load(&st->ntop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int job = blockIdx.x; job < nsub; job += gridDim.x) {
        const int4 jb = sub[job];
        const uint32_t cb0 = cbase[job], cb1 = cbase[job + 1];
        const int base = top + (int)cb0;  // ids reserved by subtree_base_kernel
        const int a = jb.x, cnt = jb.y - jb.x, L0 = jb.z, slot = jb.w;
        if (cnt < 2 || cnt > kSubtreeMax || L0 < 1 || L0 > kKeyLevels) {
            if (lane == 0) {
                atomicOr(&st->err, kErrCapacity);
                subcells[job] = make_int4(0, 0, L0, L0 - 1);
            }
            continue;
        }
        // every global load of the job is issued before the first wait: l[]
        // (lane-strided), then the lane's own chunk [i0, i1) of keys and bodies
        const int chunk = (cnt + 63) >> 6;
        const int i0 = min(cnt, lane * chunk), i1 = min(cnt, i0 + chunk);
        int lv[C];
        uint64_t kq[C];
        int pq[C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int i = lane + 64 * q;
            lv[q] = i < cnt - 1 ? (int)lcp[a + i] : L0 - 1;
        }
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const bool in = i0 + q < i1;
            kq[q] = in ? skeys[a + i0 + q] : 0ull;
            pq[q] = in ? perm[a + i0 + q] : 0;
        }
#pragma unroll
        for (int q = 0; q < C; ++q)
            if (lane + 64 * q < cnt) s_min[0][lane + 64 * q] = (uint8_t)lv[q];
        __syncthreads();
        int mine = 0, deep = L0;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int i = i0 + q;
            if (i >= i1) break;
            const int P = i == 0 ? L0 - 1 : (int)s_min[0][i - 1];
            const int l = s_min[0][i];
            mine += max(0, l - P);
            deep = max(deep, l);
        }
        int ks;
        int run = wave_exclusive_sum(mine, &ks);
        for (int o = 32; o > 0; o >>= 1) deep = max(deep, __shfl_xor(deep, o));
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int i = i0 + q;
            if (i >= i1) break;
            const int P = i == 0 ? L0 - 1 : (int)s_min[0][i - 1];
            s_cs[i] = (uint16_t)run;
            run += max(0, (int)s_min[0][i] - P);
        }
        if (ks < 1 || ks > 65535 || ks != (int)(cb1 - cb0) || (int64_t)base + ks > (int64_t)m) {
            if (lane == 0) {
                atomicOr(&st->err, kErrCapacity);
                subcells[job] = make_int4(0, 0, L0, L0 - 1);
            }
            __syncthreads();
            continue;
        }
        // range-minimum table over the cnt - 1 shared-digit counts
        for (int j = 1; j < kLiftLevels && (1 << j) <= cnt - 1; ++j) {
            __syncthreads();
            const int h = 1 << (j - 1);
            for (int x = lane; x + (1 << j) <= cnt - 1; x += 64)
                s_min[j][x] = (uint8_t)min((int)s_min[j - 1][x], (int)s_min[j - 1][x + h]);
        }
        // empty slots, cell levels and starts
        int *cslots = child + 8 * (int64_t)base;
        for (int k = lane; k < 8 * ks; k += 64) cslots[k] = -1;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int i = i0 + q;
            if (i >= i1) break;
            const int P = i == 0 ? L0 - 1 : (int)s_min[0][i - 1];
            const int l = s_min[0][i];
            for (int L = P + 1; L <= l; ++L) {
                const int c = base + (int)s_cs[i] + (L - P - 1);
                cell_level[c] = L;
                start[c] = a + i;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        // every cell (but the root) and every body into its parent's slot
        bool bad = false;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const int i = i0 + q;
            if (i >= i1) break;
            const int P = i == 0 ? L0 - 1 : (int)s_min[0][i - 1];
            const int l = s_min[0][i];
            const uint64_t key = kq[q];
            const int own = (int)s_cs[i] - P - 1;  // id(i, L) - base = own + L
            // start of the level-P run through i (only used for i >= 1)
            int s = 0;
            if (i >= 1) {
                int p = i - 1;
                for (int j = kLiftLevels - 1; j >= 0; --j) {
                    const int w = 1 << j;
                    if (p - w >= 0 && (int)s_min[j][p - w] >= P) p -= w;
                }
                s = p;
            }
            const int Ps = s == 0 ? L0 - 1 : (int)s_min[0][s - 1];
            const int up = (int)s_cs[s] - Ps - 1 + P;  // id(s, P) - base
            for (int L = max(P + 1, L0 + 1); L <= l; ++L) {
                const int par = L - 1 > P ? own + L - 1 : up;
                const int d = digit_at(key, L - 1);
                if (par < 0 || par >= ks) { bad = true; continue; }
                child[8 * ((int64_t)base + par) + d] = base + own + L;
            }
            const int D = max(P, l);
            const int par = l > P ? own + l : up;
            const int d = D < kKeyLevels ? digit_at(key, D) : (l > P ? 0 : i - s);
            if (par < 0 || par >= ks || d >= 8) { bad = true; continue; }
            child[8 * ((int64_t)base + par) + d] = pq[q];
        }
        if (__ballot(bad) && lane == 0) atomicOr(&st->err, kErrBucket);
        if (lane == 0) {
            child[slot] = base;
            subcells[job] = make_int4(base, ks, L0, deep);
        }
        // one wave per workgroup: its LDS operations are in order, so the next
        // job may overwrite the tables without waiting for these stores
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------ summary
struct CellSum {
    double ms, mx, my, mz;
    int cnt, first;
};

// Sums the children of c in slot order; a level-21 bucket first puts its
// bodies (equal keys) in ascending index order, the stable key order.  The
// children of a cell in slot order are consecutive ranges of the sorted
// bodies, so with the cell's first sorted position known (first >= 0) a body
// child is read from the sorted copy at a running cursor; otherwise (first < 0)
// from its own index, with its sorted position from pos[].
// The slots go in batches of B (a divisor of 8): per batch (1) the cell
// children's sums, counts and (without a cursor) first bodies, every load
// issued before the first use (a slot without a cell loads node 0, unused),
// (2) the body children at the cursor (the cell's first body plus the sizes of
// the slots before them), (3) the sums in slot order, in FP64.  B = 2 for the
// subtrees (more would cost occupancy: 4 needs 90 registers against 59), 4 for
// the top cells.  CURSOR: the build has given every cell its first sorted body
// (first >= 0), which the summary then neither reads for the children nor
// stores again.  LCNT: the counts of the cells [lbase, ...) of a subtree are
// in LDS at lcnt (every cell child of a subtree cell is one of them).
template <int B, bool CURSOR, bool LCNT>
__device__ inline CellSum sum_cell(int c, int first, bool bucket, int n, int *child, const int *lslots,
                                   const float4 *bodies4, const float4 *sorted4, const int32_t *pos,
                                   const double4 *sum64, const int *count, const int *start,
                                   const uint16_t *lcnt, int lbase)
{
    int v[8];
    const int *src = lslots ? lslots : child + 8 * (int64_t)c;
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = src[s];
    if (bucket) {
        for (int i = 1; i < 8; ++i) {
            int a = v[i], j = i - 1;
            while (j >= 0 && (v[j] == -1 || (a != -1 && v[j] > a))) { v[j + 1] = v[j]; --j; }
            v[j + 1] = a;
        }
        for (int s = 0; s < 8; ++s) child[8 * (int64_t)c + s] = v[s];
    }
    CellSum r{0.0, 0.0, 0.0, 0.0, 0, 0x7fffffff};
    int cur = first;
#pragma unroll
    for (int s0 = 0; s0 < 8; s0 += B) {
        double cm[B], cx[B], cy[B], cz[B];
        int kc[B], kf[B], bp[B];
        float4 b[B];
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const int64_t q = v[s0 + t] >= n ? v[s0 + t] : 0;
            const double4 cs = sum64[q];  // one 32-byte record: mass and the three moments
            cm[t] = cs.x;
            cx[t] = cs.y;
            cy[t] = cs.z;
            cz[t] = cs.w;
            kc[t] = LCNT ? (v[s0 + t] >= n ? lcnt[v[s0 + t] - lbase] : 0) : count[q];
            kf[t] = CURSOR ? 0 : start[q];
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const int ch = v[s0 + t];
            bp[t] = cur;
            if (cur >= 0) cur += ch == -1 ? 0 : (ch < n ? 1 : kc[t]);
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const int ch = v[s0 + t];
            const bool body = ch >= 0 && ch < n;
            const int p = bp[t];
            const float4 *a = body && p >= 0 && p < n ? sorted4 + p : bodies4 + (body ? ch : 0);
            b[t] = *a;
        }
#pragma unroll
        for (int t = 0; t < B; ++t) {
            const int ch = v[s0 + t];
            if (ch == -1) continue;
            if (ch < n) {
                const double mb = (double)b[t].w;
                r.ms = __dadd_rn(r.ms, mb);
                r.mx = __dadd_rn(r.mx, __dmul_rn(mb, (double)b[t].x));
                r.my = __dadd_rn(r.my, __dmul_rn(mb, (double)b[t].y));
                r.mz = __dadd_rn(r.mz, __dmul_rn(mb, (double)b[t].z));
                r.cnt += 1;
                if (!CURSOR) r.first = min(r.first, bp[t] >= 0 ? bp[t] : pos[ch]);
            } else {
                r.ms = __dadd_rn(r.ms, cm[t]);
                r.mx = __dadd_rn(r.mx, cx[t]);
                r.my = __dadd_rn(r.my, cy[t]);
                r.mz = __dadd_rn(r.mz, cz[t]);
                r.cnt += kc[t];
                if (!CURSOR) r.first = min(r.first, kf[t]);
            }
        }
    }
    return r;
}

struct SumOut {
    int n;
    int *child;
    const float4 *bodies4;   // bodies (x, y, z, m) by index
    float *mass;             // cells written (single system)
    float *cx, *cy, *cz;     // cell centre of mass (single system; nullptr for the ensemble)
    int *count, *start;      // start: set by the build for every cell when `cursor`
    const int32_t *pos, *cell_level;
    double4 *sum64;          // FP64 cell sums {mass, moments x, y, z} (written)
    float4 *nodes4;          // bodies in key order at [0, n) (read), cells at [n, m) (written)
    int cursor;              // cell starts known before the summary
};

// lev: the cell's level when the caller has it (else -1); lslots: an LDS copy of
// its slots (else nullptr); with LCNT, lcnt holds the subtree's counts in LDS
// from lbase and gets this cell's count too, and lfirst is the cell's first
// sorted body (staged by the caller; with CURSOR only).
// B: slots whose loads go together; CURSOR: o.cursor (sum_cell).
template <int B, bool CURSOR, bool LCNT = false>
__device__ inline void summarise(const SumOut &o, int c, DeviceState *st, bool single_root, int root_first = 0,
                                 int lev = -1, const int *lslots = nullptr, uint16_t *lcnt = nullptr, int lbase = 0,
                                 int lfirst = -1)
{
    const bool bucket = !single_root && (lev >= 0 ? lev : o.cell_level[c]) == kKeyLevels;
    const int first = !CURSOR ? -1 : single_root ? root_first : LCNT ? lfirst : o.start[c];
    const CellSum r = sum_cell<B, CURSOR, LCNT>(c, first, bucket, o.n, o.child, lslots, o.bodies4, o.nodes4, o.pos, o.sum64,
                                  o.count, o.start, lcnt, lbase);
    if (r.cnt > 0 && !(r.ms > 0.0)) atomicOr(&st->err, kErrNonFinite);
    const double gx = r.mx / r.ms, gy = r.my / r.ms, gz = r.mz / r.ms;
    const float fx = (float)gx, fy = (float)gy, fz = (float)gz, fm = (float)r.ms;
    if (single_root) {
        st->root64[0] = r.ms; st->root64[1] = gx; st->root64[2] = gy; st->root64[3] = gz;
        st->root4 = make_float4(fx, fy, fz, fm);
        return;
    }
    o.sum64[c] = make_double4(r.ms, r.mx, r.my, r.mz);
    if (o.cx) {
        o.cx[c] = fx; o.cy[c] = fy; o.cz[c] = fz; o.mass[c] = fm;
    }
    o.count[c] = r.cnt;
    if (LCNT) lcnt[c - lbase] = (uint16_t)r.cnt;  // a subtree holds at most kSubtreeMax bodies
    if (!CURSOR) o.start[c] = r.first;
    o.nodes4[c] = make_float4(fx, fy, fz, fm);
}

// Subtree summary: workgroup w sums subtrees w, w + gridDim.x, ...: with as many
// workgroups as subtrees, about, the dispatcher hands each free slot of a CU
// the next workgroup, a job queue with no shared counter.  The cells of a
// subtree are the contiguous ids [base, base + ks), summed deepest
// level first with a barrier between levels.  A subtree of up to kSumLdsCells
// cells first copies its cells' levels and slots into LDS (one coalesced read),
// and keeps the count of every cell it has summed there, so a level costs the
// reads of the children's FP64 sums and bodies and the stores only.
constexpr int kSumLdsCells = 512;
constexpr int kSubSumBatch = 2;
template <bool CURSOR>
__global__ void __launch_bounds__(256) subtree_summary_kernel(SumOut o, const int4 *__restrict__ subcells,
                                                              DeviceState *st)
{
    __shared__ int s_job;
    // 19,972 bytes of LDS (20 KB as allocated): eight workgroups of 256 threads
    // fit a CU's 160 KB
    __shared__ uint8_t s_lev[kSumLdsCells];
    __shared__ int s_slot[8 * kSumLdsCells];
    __shared__ uint16_t s_cnt[kSumLdsCells];  // counts of the subtree's summed cells
    __shared__ int s_first[kSumLdsCells];     // first sorted body of each cell (CURSOR)
    static_assert(kSubtreeMax < 65536 && kKeyLevels < 256, "counts and levels fit the LDS types");
    const int tid = threadIdx.x;
    const int nsub = __hip_atomic_load(&st->nsub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int next = blockIdx.x;
    for (;;) {
        // One read of the latch per job, by thread 0, published with the job:
        // the workgroup continues or leaves as a whole (a zero cell mass raises
        // the latch from inside this launch).
        if (tid == 0) s_job = latched(st) ? nsub : next;
        next += gridDim.x;
        __syncthreads();
        const int job = s_job;
        __syncthreads();
        if (job >= nsub) break;
        const int4 sc = subcells[job];
        const int base = sc.x, ks = sc.y, L0 = sc.z;
        const bool lds = ks <= kSumLdsCells;
        if (lds) {
            for (int k = tid; k < ks; k += blockDim.x) {
                s_lev[k] = (uint8_t)o.cell_level[base + k];
                if (CURSOR) s_first[k] = o.start[base + k];
            }
            for (int k = tid; k < 8 * ks; k += blockDim.x) s_slot[k] = o.child[8 * (int64_t)base + k];
            __syncthreads();
        }
        if (lds) {  // levels, slots and the counts of summed cells in LDS
            for (int L = sc.w; L >= L0; --L) {
                for (int k = tid; k < ks; k += blockDim.x)
                    if (s_lev[k] == L)
                        summarise<kSubSumBatch, CURSOR, true>(o, base + k, st, false, 0, L, s_slot + 8 * k, s_cnt,
                                                              base, CURSOR ? s_first[k] : -1);
                __syncthreads();
            }
        } else {
            for (int L = sc.w; L >= L0; --L) {
                for (int k = tid; k < ks; k += blockDim.x)
                    if (o.cell_level[base + k] == L) summarise<kSubSumBatch, CURSOR>(o, base + k, st, false, 0, L);
                __syncthreads();
            }
        }
    }
}

// Top summary: the top cells level by level, deepest first, then the roots.  A
// level's cells read only the level below.  A cell's sum gathers from up to 8
// children in unrelated cache lines, and one CU's address pipeline takes such
// gathers a few lanes at a time, so a level of more than kTopNarrow cells is
// spread over kTopBlocks workgroups (its cell i to workgroup i mod kTopBlocks)
// that meet at a grid barrier after it; the narrower levels above it (a top cell's
// parent is a top cell: a level never has more cells than the one below) and
// the roots are summed by workgroup 0 alone, with workgroup barriers.
// kTopBlocks workgroups of 256 threads fit on the device at once, so every
// workgroup reaches every barrier; the wait is bounded all the same: one that
// runs out latches kErrStall, and every wave still finishes.
constexpr int kTopBlocks = 64;
constexpr int kTopNarrow = 64;
constexpr int kTopSpin = 1 << 20;  // polls of one barrier wait (about a second)

// Grid barrier number `phase` (1, 2, ...) of the launch.  Every thread makes its
// stores visible device-wide (release), thread 0 arrives and waits until every
// workgroup has, and every thread then acquires what they wrote.
__device__ inline void top_barrier(DeviceState *st, int phase, bool &stalled)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const int target = phase * (int)gridDim.x;
        __hip_atomic_fetch_add(&st->tsync, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; !stalled && __hip_atomic_load(&st->tsync, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target;
             ++k) {
            if (k == kTopSpin) {
                atomicOr(&st->err, kErrStall);
                stalled = true;  // no further waits in this workgroup
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <bool CURSOR>
__global__ void __launch_bounds__(256) top_summary_kernel(SumOut o, Segs g, int base0, int nsamples,
                                                          const uint32_t *__restrict__ top_id, DeviceState *st)
{
    __shared__ int s_first[kLevels + 1];
    __shared__ int s_stop;
    const int tid = threadIdx.x;
    if (tid == 0) s_stop = latched(st);  // one read for the whole workgroup
    // The top cells of level L are the ids base0 + [s_first[L], s_first[L + 1])
    // (the scanned, level-major top table); every workgroup reads the same.
    if (tid <= kLevels)
        s_first[tid] = tid < kLevels ? (int)top_id[tid * nsamples]
                                     : __hip_atomic_load(&st->ntop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // Latched before this launch: every workgroup reads it so and skips the sums,
    // but still meets every barrier (a latch raised by this launch's own sums may
    // be seen by some workgroups only).
    const bool stop = s_stop != 0;
    bool stalled = false;
    int L = kKeyLevels, phase = 0;
    for (; L >= 1; --L) {
        const int cnt = s_first[L + 1] - s_first[L];
        if (cnt <= 0) continue;  // uniform: no barrier for an empty level
        if (cnt <= kTopNarrow) break;
        if (!stop)
            for (int i = blockIdx.x + gridDim.x * tid; i < cnt; i += gridDim.x * blockDim.x)
                summarise<4, CURSOR>(o, base0 + s_first[L] + i, st, false, 0, L);
        top_barrier(st, ++phase, stalled);
    }
    if (blockIdx.x == 0 && !stop) {
        for (; L >= 1; --L) {
            const int cnt = s_first[L + 1] - s_first[L];
            if (cnt <= 0) continue;
            for (int i = tid; i < cnt; i += blockDim.x) summarise<4, CURSOR>(o, base0 + s_first[L] + i, st, false, 0, L);
            __syncthreads();
        }
        for (int s = tid; s < g.nseg; s += blockDim.x) summarise<4, CURSOR>(o, seg_root(g, s), st, !g.batch);
    }
    if (tid == 0) {  // the last workgroup out (every other one is past its last wait) resets the barrier
        const int out = __hip_atomic_fetch_add(&st->tdone, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (out == (int)gridDim.x - 1) {
            __hip_atomic_store(&st->tsync, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->tdone, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ------------------------------------------------------------------ layout
// Bodies in key order as float4 (x, y, z, m) at nodes4[0..n), made right after
// the build for the summary; sorted[] (the reference's output of the sort
// stage) in the sort stage; then the frame blocks of the force walk
// (frame_layout_kernel).  The bodies are read in input order and stored at
// their sorted positions pos[] (the sort's inverse permutation): random 16-byte
// reads each fetch a whole line, random 16-byte stores only dirty their bytes.
__global__ void __launch_bounds__(256) sorted_bodies_kernel(int n, const int32_t *__restrict__ pos,
                                                            const float4 *__restrict__ bodies4,
                                                            float4 *__restrict__ nodes4, const DeviceState *st)
{
    if (latched(st)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) nodes4[pos[i]] = bodies4[i];
}

__global__ void __launch_bounds__(256) sorted_index_kernel(int n, const int32_t *__restrict__ perm,
                                                           int *__restrict__ sorted, const DeviceState *st)
{
    if (latched(st)) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) sorted[k] = perm[k];
}

// Frame blocks.  Block b lists the non-empty children of one cell, compacted in
// slot order, as the records [off(b), off(b) + k_b) of the packed arrays:
// fnode[off(b) + k] = (x, y, z, m) of the child and fmeta[off(b) + k] its
// descriptor {tag, crit}:
//   body:  tag = sorted position << 4,        crit = -1 (always accepted);
//   cell:  tag = off(block) << 4 | children (1-8), crit = crit_L + eps^2,
// crit_L = (l_L/theta)^2 computed as in the opening test (crit_0 rounded once
// from FP64, scaled by 4^-L exactly) and eps^2 added once in FP32, so the walk
// compares crit < |d|^2 + eps^2 -- the value its interaction needs anyway.
// A cell child is described by the bottom of its chain, the first descendant
// with two or more children: the cells of a one-child chain carry the same mass
// and centre of mass (the FP64 sums of a single child are exact copies), and
// crit_L + eps^2 falls with L, so a body accepts some cell of the chain exactly
// when it accepts the bottom one, and meets the same monopole once either way.  fnode
// keeps the data of the child itself.
// Single system: block 0 is the root (slots child[0..7]), block i >= 1 is cell
// n + i - 1.  Ensemble: block i is cell n + i, the roots being blocks 0..S-1.
// froot4 / frootm describe each root as a child of a virtual parent.  The
// parent blocks of every block and of every body (by sorted position) go to
// pblk / bpar for the adjoint's top-down pass.
// Blocks follow in id order with no gaps between them, so off(b) is the number
// of children of the blocks before b: the packed records take 24 bytes per node
// (fnode and fmeta) where blocks of eight slots took 192 bytes per cell, and a
// frame's three block loads come from fewer distinct lines of the L2.
template <bool BATCH>
__device__ inline int block_of(int c, int n)
{
    return BATCH ? c - n : (c == 0 ? 0 : c - n + 1);
}

// Children per block, summed within each run of kOffRun blocks: boff[b] = the
// children of the blocks of b's run before b, bsum[run] = the run's total (the
// runs are then scanned in place).  One workgroup per run.  The slots are read
// as 16-byte halves of the blocks' rows, consecutive threads on consecutive
// halves, and the two halves of a block meet by a lane exchange.
template <bool BATCH>
__global__ void __launch_bounds__(256) frame_count_kernel(int n, const int *__restrict__ index,
                                                          const int *__restrict__ child, uint32_t *__restrict__ boff,
                                                          uint32_t *__restrict__ bsum, const DeviceState *st)
{
    __shared__ uint32_t cnt[kOffRun];
    __shared__ uint32_t sh[4];
    static_assert(kOffRun == 4 * 256, "four blocks per thread");
    if (latched(st)) return;
    const int64_t items = BATCH ? (int64_t)*index - n : (int64_t)*index - n + 1;
    const int64_t r0 = (int64_t)blockIdx.x * kOffRun;
    if (r0 >= items) {  // uniform: a run past the last block
        if (threadIdx.x == 0) bsum[blockIdx.x] = 0u;
        return;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int h = (int)threadIdx.x + 256 * u;  // half h & 1 of the run's block h >> 1
        const int64_t item = r0 + (h >> 1);
        int k = 0;
        if (item < items) {
            const int node = BATCH ? n + (int)item : (item == 0 ? 0 : n + (int)item - 1);
            const int4 q = reinterpret_cast<const int4 *>(child + 8 * (int64_t)node)[h & 1];
            k = (q.x != -1) + (q.y != -1) + (q.z != -1) + (q.w != -1);
        }
        k += __shfl_xor(k, 1);
        if ((h & 1) == 0) cnt[h >> 1] = (uint32_t)k;
    }
    __syncthreads();
    const int j0 = 4 * (int)threadIdx.x;
    uint32_t k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u] = cnt[j0 + u];
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(k[0] + k[1] + k[2] + k[3], sh, &tot);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (r0 + j0 + u < items) boff[r0 + j0 + u] = ex;
        ex += k[u];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Opening threshold of a level-L cell as the walk compares it: crit_L + eps^2,
// crit_L = (l_L/theta)^2, l_0 the side of the root cell of the key grid.
__device__ inline float crit_of(double span, float theta, float eps2, int lev)
{
    const double lc0 = ldexp(span / kMaskCoord, kKeyLevels);
    const double r0 = lc0 / (double)theta;
    return ldexpf((float)(r0 * r0), -2 * lev) + eps2;
}

// Descriptor of the bottom of the one-child chain starting at cell c (c = 0 is
// the single-system root).
template <bool BATCH>
__device__ inline int2 chain_bottom(int c, int n, const int *__restrict__ child,
                                    const int32_t *__restrict__ cell_level, bool root, double span, float theta,
                                    float eps2, const uint32_t *__restrict__ boff, const uint32_t *__restrict__ bsum)
{
    for (;;) {
        const int4 *p = reinterpret_cast<const int4 *>(child + 8 * (int64_t)c);
        const int4 a = p[0], b = p[1];
        const int v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        int cnt = 0, last = -1;
        for (int s = 0; s < 8; ++s)
            if (v[s] != -1) { ++cnt; last = v[s]; }
        if (cnt == 1 && last >= n) {
            c = last;
            root = false;
            continue;
        }
        const int lev = root ? 0 : cell_level[c];
        const int off = frame_off(boff, bsum, block_of<BATCH>(c, n));
        return make_int2((off << 4) | cnt, __float_as_int(crit_of(span, theta, eps2, lev)));
    }
}

// chain_bottom of a cell child whose slots (a, b) and level are already loaded:
// the loads of a chain are made only when the child has a single cell child.
template <bool BATCH>
__device__ inline int2 chain_bottom_from(int c, int4 a, int4 b, int lev, int n, const int *__restrict__ child,
                                         const int32_t *__restrict__ cell_level, double span, float theta,
                                         float eps2, const uint32_t *__restrict__ boff,
                                         const uint32_t *__restrict__ bsum)
{
    for (;;) {
        const int v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        int cnt = 0, last = -1;
        for (int s = 0; s < 8; ++s)
            if (v[s] != -1) { ++cnt; last = v[s]; }
        if (cnt == 1 && last >= n) {
            c = last;
            const int4 *p = reinterpret_cast<const int4 *>(child + 8 * (int64_t)c);
            a = p[0];
            b = p[1];
            lev = cell_level[c];
            continue;
        }
        const int off = frame_off(boff, bsum, block_of<BATCH>(c, n));
        return make_int2((off << 4) | cnt, __float_as_int(crit_of(span, theta, eps2, lev)));
    }
}

// Ensemble: the system of a cell from its first sorted body, whose group key
// (made before the layout) holds the system index above the Hilbert digits.
struct SysInfo {
    const uint64_t *gkey;
    const float *theta;
    const double *ext;
};

template <bool BATCH>
__global__ void __launch_bounds__(256) frame_layout_kernel(
    int n, int nroot, const int *__restrict__ index, const int *__restrict__ child, const int32_t *__restrict__ pos,
    const int32_t *__restrict__ cell_level, const int *__restrict__ count, const int *__restrict__ start,
    const float4 *__restrict__ nodes4, float4 *__restrict__ fnode, int2 *__restrict__ fmeta,
    const uint32_t *__restrict__ boff, const uint32_t *__restrict__ bsum, float4 *__restrict__ froot4,
    int2 *__restrict__ frootm, int32_t *__restrict__ pblk, int32_t *__restrict__ bpar, float theta, float eps2,
    SysInfo si, const DeviceState *st)
{
    if (latched(st)) return;
    const int ncell = *index - n;
    const int64_t items = BATCH ? ncell : ncell + 1;
    const int64_t total = 8 * items;
    const int64_t bound = (total + 63) & ~(int64_t)63;  // whole waves iterate together (ballot below)
    const int lane = threadIdx.x & 63;
    const int body_crit = __float_as_int(-1.f);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < bound;
         k += (int64_t)blockDim.x * gridDim.x) {
        const int64_t item = k >> 3;
        const int s = (int)(k & 7);
        const int node = BATCH ? n + (int)item : (item == 0 ? 0 : n + (int)item - 1);
        const int ch = k < total ? child[8 * (int64_t)node + s] : -1;
        // Everything a cell child needs (its slots, level, count and centre) and
        // the node's first sorted body are loaded now, before the scan below: a
        // slot that holds no cell reads node 0, unused.
        const int64_t cq = ch >= n ? ch : 0;
        const int4 *cs4 = reinterpret_cast<const int4 *>(child + 8 * cq);
        const int4 ca = cs4[0], cb = cs4[1];
        const int clev = cell_level[cq];
        const int ccnt = count[cq];
        const float4 cnode = nodes4[cq];
        const int nstart = k >= total || (!BATCH && item == 0) ? 0 : start[node];  // the single root is node 0
        const int boff0 = k < total ? frame_off(boff, bsum, (int)item) : 0;  // the block's first record
        const uint64_t occ = __ballot(ch != -1);
        const unsigned seg = (unsigned)(occ >> (lane & ~7)) & 0xffu;
        const int rank = __popc(seg & ((1u << s) - 1u));
        // Sorted position of a body child: the node's first sorted body plus the
        // bodies of the slots before it (children in slot order are consecutive
        // ranges of the key order).  8-lane exclusive sum of the children's sizes.
        const int w = ch == -1 ? 0 : ch < n ? 1 : ccnt;
        int off = w;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            const int y = __shfl_up(off, o, 8);
            if (s >= o) off += y;
        }
        off -= w;
        if (ch == -1 && !(k < total && item < nroot && s == 0)) continue;
        double span;
        float th;
        if (BATCH) {
            const int sys = item < nroot ? (int)item : (int)(si.gkey[nstart] >> (3 * kGroupLevels));
            span = si.ext[4 * sys + 3];
            th = si.theta[sys];
        } else {
            span = st->ext[3];
            th = theta;
        }
        if (k < total && item < nroot && s == 0) {
            froot4[item] = BATCH ? nodes4[node] : st->root4;
            frootm[item] = chain_bottom<BATCH>(node, n, child, cell_level, true, span, th, eps2, boff, bsum);
            pblk[item] = -1;
        }
        if (ch == -1) continue;
        const int64_t slot = (int64_t)boff0 + rank;
        if (ch < n) {
            const int q = nstart + off;
            const int ps = q >= 0 && q < n ? q : pos[ch];
            fnode[slot] = nodes4[ps];
            fmeta[slot] = make_int2(ps << 4, body_crit);
            bpar[ps] = (int)item;
        } else {
            fnode[slot] = cnode;
            fmeta[slot] = chain_bottom_from<BATCH>(ch, ca, cb, clev, n, child, cell_level, span, th, eps2, boff, bsum);
            pblk[block_of<BATCH>(ch, n)] = (int)item;
        }
    }
}

// Group order of the force walk.  The walk gives each wave 64 consecutive
// bodies of this order.  Key order has jumps (consecutive Morton keys can lie in
// cells far apart), so bodies are ordered by the top kGroupLevels Hilbert digits
// of their key instead, stably over key order (ties keep key order); the
// ensemble puts the system index above them.  The order changes which bodies
// share a wave, never a body's interaction list.  A single system's keys
// (3 kGroupLevels = 30 bits) are sorted as 32-bit keys.
template <class K>
__global__ void __launch_bounds__(256) group_keys_kernel(int n, const uint64_t *__restrict__ skeys,
                                                         const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ sys_of,
                                                         K *__restrict__ gkey, const DeviceState *st)
{
    __shared__ uint8_t s_tab[kHilbertTable];
    hilbert_table_fill(s_tab);
    __syncthreads();
    if (latched(st)) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) {
        uint64_t h = hilbert_top_lds(skeys[k] & ~(1ull << 63), kGroupLevels, s_tab);
        if (sys_of) h |= (uint64_t)sys_of[perm[k]] << (3 * kGroupLevels);
        gkey[k] = (K)h;
    }
}

// ------------------------------------------------------------------ launchers
// Radix passes of the group-order keys (Hilbert digits, then the system index).
static int group_passes(int sys_bits) { return (3 * kGroupLevels + sys_bits + kRadixBits - 1) / kRadixBits; }

static inline int capped_grid(int64_t work, int cap = 65536)
{
    int64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}


// Workgroups of a job loop (subtree build, locked insertion, subtree summary):
// about one per subtree (a subtree averages about 256 bodies), at least
// kBuildBlocks.  With many jobs per workgroup the stride balances them worse:
// at N = 48.8M, 16384 summary workgroups took 3.50 ms against 3.30 ms for a
// shared job counter.
static int job_grid(int nbodies)
{
    return std::min(1 << 20, std::max(kBuildBlocks, nbodies / 256));
}

static void build_tree(Segs g, int nsamples, int base0, int m, TreeScratch &ts, DeviceState *st, hipStream_t s,
                       int n, const uint64_t *skeys, const int32_t *perm, const uint64_t *keys, int *child,
                       int *index, int32_t *cell_level, int *start, bool locked)
{
    lcp_kernel<<<capped_grid(n), 256, 0, s>>>(n, skeys, ts.lcp, st);
    top_cells_kernel<<<(32 * nsamples + 255) / 256, 256, 0, s>>>(g, nsamples, skeys, ts.top_flag, ts.top_id, ts.top_a,
                                                             ts.top_b, st);
    launch_exclusive_scan_u32(ts.top_id, nsamples * kLevels, ts.scan_tmp, s);
    const int wire = nsamples * kLevels + g.nseg;  // entries, 8 lanes each
    top_wire_kernel<<<(8 * wire + 255) / 256, 256, 0, s>>>(g, nsamples, skeys, perm, child, cell_level, ts.top_flag,
                                                        ts.top_id, ts.top_a, ts.top_b, ts.sub, ts.sub_max,
                                                        base0, m, index, start, st);
    subtree_cells_kernel<<<capped_grid(64 * (int64_t)ts.sub_max, 4096), 256, 0, s>>>(ts.sub, ts.sub_max, ts.lcp,
                                                                                     ts.cbase, st);
    subtree_base_kernel<<<1, 1024, 0, s>>>(ts.cbase, ts.sub_max, base0, m, index, st);
    if (locked)
        subtree_insert_kernel<<<job_grid(g.n), 256, 0, s>>>(n, m, skeys, perm, keys, ts.lcp, child, base0,
                                                                ts.cbase, cell_level, ts.sub, ts.sub_max,
                                                                ts.subcells, st);
    else
        subtree_build_kernel<<<job_grid(g.n), 64, 0, s>>>(m, skeys, perm, ts.lcp, child, base0, ts.cbase, cell_level,
                                                         start, ts.sub, ts.sub_max, ts.subcells, st);
}

// base0: the first top cell's id; nsamples: the top table's samples.
static void summarise_tree(Segs g, int base0, int nsamples, TreeScratch &ts, DeviceState *st, hipStream_t s,
                           const SumOut &o)
{
    if (o.cursor) {
        subtree_summary_kernel<true><<<job_grid(g.n), 256, 0, s>>>(o, ts.subcells, st);
        top_summary_kernel<true><<<kTopBlocks, 256, 0, s>>>(o, g, base0, nsamples, ts.top_id, st);
    } else {
        subtree_summary_kernel<false><<<job_grid(g.n), 256, 0, s>>>(o, ts.subcells, st);
        top_summary_kernel<false><<<kTopBlocks, 256, 0, s>>>(o, g, base0, nsamples, ts.top_id, st);
    }
}

static inline int samples_of(int n) { return (n + kSubtreeMax - 1) / kSubtreeMax; }

bool alloc_tree_scratch(TreeScratch &ts, int n, int nseg, int samples)
{
    ts.samples_max = samples;
    int64_t sub = 8 * ((int64_t)samples * kKeyLevels + nseg);
    sub = std::min<int64_t>(sub, (int64_t)n / 2 + 8 * (int64_t)nseg + 8);
    ts.sub_max = (int)sub;
    const size_t tt = (size_t)samples * kLevels;
    return hipMalloc((void **)&ts.lcp, (size_t)n + 1) == hipSuccess &&
           hipMalloc((void **)&ts.top_flag, 4 * tt) == hipSuccess &&
           hipMalloc((void **)&ts.top_id, 4 * tt) == hipSuccess &&
           hipMalloc((void **)&ts.top_a, 4 * tt) == hipSuccess &&
           hipMalloc((void **)&ts.top_b, 4 * tt) == hipSuccess &&
           hipMalloc((void **)&ts.sub, sizeof(int4) * (size_t)ts.sub_max) == hipSuccess &&
           hipMalloc((void **)&ts.subcells, sizeof(int4) * (size_t)ts.sub_max) == hipSuccess &&
           hipMalloc((void **)&ts.cbase, sizeof(uint32_t) * ((size_t)ts.sub_max + 1)) == hipSuccess &&
           hipMalloc((void **)&ts.scan_tmp, 4 * 1024) == hipSuccess && (int64_t)samples * kLevels <= (1 << 20);
}

void free_tree_scratch(TreeScratch &ts)
{
    void *p[] = {ts.lcp, ts.top_flag, ts.top_id, ts.top_a, ts.top_b, ts.sub, ts.subcells, ts.scan_tmp,
                 ts.cbase};
    for (void *q : p)
        if (q) hipFree(q);
    ts = TreeScratch{};
}

void launch_build(nbx_context *c, hipStream_t s, int n, int m, int *child, int *index)
{
    if (n <= 0) return;
    Segs g{nullptr, nullptr, 1, n, 0};
    build_tree(g, samples_of(n), n, m, c->ts, c->st, s, n, c->skeys, c->perm, c->keys, child, index, c->cell_level,
               c->start, c->p.build_locked != 0);
    sorted_bodies_kernel<<<capped_grid(n), 256, 0, s>>>(n, c->pos, c->bodies4, c->nodes4, c->st);
}

void launch_summarize(nbx_context *c, hipStream_t s)
{
    Segs g{nullptr, nullptr, 1, c->n, 0};
    SumOut o{c->n, c->child, c->bodies4, c->mass, c->x, c->y, c->z, c->count, c->start, c->pos,
             c->cell_level, c->sum64, c->nodes4, c->p.build_locked == 0};
    summarise_tree(g, c->n, samples_of(c->n), c->ts, c->st, s, o);
}

void launch_layout(nbx_context *c, hipStream_t s, int *sorted)
{
    const int n = c->n;
    sorted_index_kernel<<<capped_grid(n), 256, 0, s>>>(n, c->perm, sorted, c->st);
    launch_frame_layout(c, s);
    static_assert(3 * kGroupLevels <= 32, "single-system group keys are 32-bit");
    group_keys_kernel<uint32_t><<<capped_grid(n), 256, 0, s>>>(n, c->skeys, c->perm, nullptr, c->gkey, c->st);
    launch_radix_passes(c->hist, s, c->gkey, nullptr, n, 0, group_passes(0), c->gkey2, c->gorder,
                        reinterpret_cast<uint32_t *>(c->kbuf), c->vbuf, nullptr);
}

void launch_frame_layout(nbx_context *c, hipStream_t s)
{
    SysInfo si{};
    const int runs = (int)(((int64_t)c->m - c->n + 1 + kOffRun - 1) / kOffRun);
    frame_count_kernel<false><<<runs, 256, 0, s>>>(c->n, c->index, c->child, c->boff, c->bsum, c->st);
    launch_scan_one_block(c->bsum, runs, s);
    frame_layout_kernel<false><<<capped_grid(8 * (int64_t)(c->m - c->n + 1)), 256, 0, s>>>(
        c->n, 1, c->index, c->child, c->pos, c->cell_level, c->count, c->start, c->nodes4, c->fnode, c->fmeta, c->boff,
        c->bsum, c->froot4, c->frootm, c->pblk, c->bpar, c->p.theta, layout_eps2(c->p), si, c->st);
    c->layout_theta = c->p.theta;
    c->layout_eps2 = layout_eps2(c->p);
}

void launch_batch_build(nbx_batch *b, hipStream_t s)
{
    Segs g{b->off, b->samp_off, b->S, b->N, 1};
    build_tree(g, b->ts.samples_max, b->N + b->S, b->m, b->ts, b->st, s, b->N, b->skeys, b->perm, b->keys, b->child,
               b->index, b->cell_level, b->start, b->p.build_locked != 0);
    sorted_bodies_kernel<<<capped_grid(b->N), 256, 0, s>>>(b->N, b->pos, b->bodies4, b->nodes4, b->st);
}

void launch_batch_summarize(nbx_batch *b, hipStream_t s)
{
    Segs g{b->off, b->samp_off, b->S, b->N, 1};
    SumOut o{b->N, b->child, b->bodies4, b->mass, nullptr, nullptr, nullptr, b->count, b->start, b->pos,
             b->cell_level, b->sum64, b->nodes4, b->p.build_locked == 0};
    summarise_tree(g, b->N + b->S, b->ts.samples_max, b->ts, b->st, s, o);
}

void launch_batch_layout(nbx_batch *b, hipStream_t s)
{
    sorted_index_kernel<<<capped_grid(b->N), 256, 0, s>>>(b->N, b->perm, b->sorted, b->st);
    group_keys_kernel<uint64_t><<<capped_grid(b->N), 256, 0, s>>>(b->N, b->skeys, b->perm, b->sys_of, b->gkey, b->st);
    SysInfo si{b->gkey, b->theta, b->ext};
    const int runs = (int)(((int64_t)b->m - b->N + kOffRun - 1) / kOffRun);
    frame_count_kernel<true><<<runs, 256, 0, s>>>(b->N, b->index, b->child, b->boff, b->bsum, b->st);
    launch_scan_one_block(b->bsum, runs, s);
    frame_layout_kernel<true><<<capped_grid(8 * (int64_t)(b->m - b->N)), 256, 0, s>>>(
        b->N, b->S, b->index, b->child, b->pos, b->cell_level, b->count, b->start, b->nodes4, b->fnode, b->fmeta, b->boff,
        b->bsum, b->froot4, b->frootm, b->pblk, b->bpar, 0.f, layout_eps2(b->p), si, b->st);
    int sys_bits = 0;
    while ((1 << sys_bits) < b->S) ++sys_bits;
    launch_radix_passes(b->hist, s, b->gkey, nullptr, b->N, 0, group_passes(sys_bits), b->gkey2, b->gorder, b->kbuf,
                        b->vbuf, nullptr);
}

}  // namespace nbx
