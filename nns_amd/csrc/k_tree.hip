// k_tree.hip -- tree construction (atomic child-slot locking), bottom-up summary,
// and the traversal layout.
//
// Cells are prefixes of the 63-bit Morton key.  A cell at level L selects the
// slot of a body by key digit L (bits 62-3L .. 60-3L), so slot order is key
// order.  Single system: the root is node 0 with its slots at child[0..7] (the
// reference's convention, src/Kernels.cu:195,225) and cells take ids n, n+1, ...
// from *index.  Ensemble: system s has the root cell N + s, other cells follow.
#include "nbx_device.h"

namespace nbx {

// Lock-based insertion of one body per thread, bodies taken in key order.
//
// Each iteration a lane reads the slot for its body at the current cell
// (relaxed agent-scope atomic loads; L2 is the coherence point):
//   cell id (>= n)  descend;
//   -1              CAS -1 -> body, done;
//   -2              locked by another lane, retry;
//   body            CAS body -> -2 (lock); the winner creates the chain of
//                   cells down to the first level where the two keys differ
//                   (a level-21 bucket if the keys are equal), fences, and
//                   publishes the first cell with an atomic store (unlock).
// No lane ever waits inside an iteration on another lane, so a wave whose
// lanes contend for one slot makes progress: the CAS winner finishes its
// insertion in the same iteration.  Cell ids for a wave's splits come from
// one atomicAdd on *index (wave prefix sum).
template <bool BATCH>
__global__ void __launch_bounds__(256) build_kernel(
    int n, int m, const uint64_t *__restrict__ skeys, const int32_t *__restrict__ perm,
    const uint64_t *__restrict__ keys, int *child, int *index, int32_t *__restrict__ cell_level,
    int32_t *__restrict__ parent, int32_t *__restrict__ arrive, DeviceState *st,
    const int32_t *__restrict__ sys_of, int rootbase)
{
    if (latched(st)) return;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool done = p >= n;
    const int b = done ? 0 : perm[p];
    const uint64_t key = done ? 0ull : skeys[p];
    int node = (BATCH && !done) ? rootbase + sys_of[b] : 0;
    int L = 0;
    for (;;) {
        int need = 0, addr = 0, found = 0, ld = 0;
        uint64_t key2 = 0;
        if (!done) {
            if (L == kKeyLevels) {
                // bucket cell: first free slot
                bool placed = false, full = true;
                for (int s = 0; s < 8 && !placed; ++s) {
                    int ch = ld_slot(&child[8 * (int64_t)node + s]);
                    if (ch == -1) {
                        full = false;
                        if (cas_slot(&child[8 * (int64_t)node + s], -1, b) == -1) placed = true;
                    }
                }
                if (placed) {
                    done = true;
                } else if (full) {
                    atomicOr(&st->err, kErrBucket);
                    done = true;
                }
            } else {
                addr = 8 * node + digit_at(key, L);
                const int ch = ld_slot(&child[addr]);
                if (ch >= n) {
                    node = ch;
                    ++L;
                } else if (ch == -1) {
                    if (cas_slot(&child[addr], -1, b) == -1) done = true;
                } else if (ch >= 0) {
                    if (cas_slot(&child[addr], ch, -2) == ch) {
                        found = ch;
                        key2 = keys[ch];
                        ld = L + 1;
                        while (ld < kKeyLevels && digit_at(key, ld) == digit_at(key2, ld)) ++ld;
                        need = ld - L;  // cells at levels L+1 .. ld
                    }
                }
            }
        }
        const uint64_t wants = __ballot(need > 0);
        if (wants) {
            // inclusive prefix sum of `need` over the wave
            int incl = need;
            for (int o = 1; o < 64; o <<= 1) {
                int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const int total = __shfl(incl, 63);
            const int leader = __ffsll((unsigned long long)wants) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(index, total);
            base = __shfl(base, leader);
            if (need > 0) {
                const int first = base + incl - need;
                if ((int64_t)first + need > (int64_t)m) {
                    atomicOr(&st->err, kErrCapacity);
                    st_slot(&child[addr], found);  // unlock, tree stays consistent
                    done = true;
                } else {
                    const int4 neg = make_int4(-1, -1, -1, -1);
                    for (int j = 0; j < need; ++j) {
                        const int c = first + j;
                        const int lvl = L + 1 + j;
                        int4 *c4 = reinterpret_cast<int4 *>(child + 8 * (int64_t)c);
                        c4[0] = neg;
                        c4[1] = neg;
                        cell_level[c] = lvl;
                        parent[c] = (j == 0) ? node : c - 1;
                        arrive[c] = 0;
                        if (lvl < ld) {
                            child[8 * c + digit_at(key, lvl)] = c + 1;
                        } else if (ld < kKeyLevels) {
                            child[8 * c + digit_at(key2, ld)] = found;
                            child[8 * c + digit_at(key, ld)] = b;
                        } else {
                            child[8 * c + 0] = found;  // bucket; put in index order later
                            child[8 * c + 1] = b;
                        }
                    }
                    __threadfence();
                    st_slot(&child[addr], first);  // publish + unlock
                    done = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
    }
}

// Bottom-up summary.  A cell without cell children is summed by its own
// thread; a cell with k cell children is summed by the thread whose child
// arrives k-th (atomic counter, release fence before the increment, acquire
// fence after).  Every cell is therefore summed by exactly one thread, over its
// children in slot order, in FP64: results do not depend on scheduling.
__device__ inline int count_cell_children(const int *child, int node, int n)
{
    int k = 0;
    for (int s = 0; s < 8; ++s) k += child[8 * (int64_t)node + s] >= n;
    return k;
}

struct CellSum {
    double ms, mx, my, mz;
    int cnt, first;
};

// Sums the children of c in slot order; a level-21 bucket first puts its
// bodies (equal keys) in ascending index order, the stable key order.
__device__ inline CellSum sum_cell(int c, bool bucket, int n, int *child, const float *x, const float *y,
                                   const float *z, const float *mass, const int32_t *pos, const double *mass64,
                                   const double *mom64, const int *count, const int *start)
{
    int *slots = child + 8 * (int64_t)c;
    if (bucket) {
        int v[8];
        for (int s = 0; s < 8; ++s) v[s] = slots[s];
        for (int i = 1; i < 8; ++i) {
            int a = v[i], j = i - 1;
            while (j >= 0 && (v[j] == -1 || (a != -1 && v[j] > a))) { v[j + 1] = v[j]; --j; }
            v[j + 1] = a;
        }
        for (int s = 0; s < 8; ++s) slots[s] = v[s];
    }
    CellSum r{0.0, 0.0, 0.0, 0.0, 0, 0x7fffffff};
    for (int s = 0; s < 8; ++s) {
        const int ch = slots[s];
        if (ch == -1) continue;
        if (ch < n) {
            const double mb = (double)mass[ch];
            r.ms = __dadd_rn(r.ms, mb);
            r.mx = __dadd_rn(r.mx, __dmul_rn(mb, (double)x[ch]));
            r.my = __dadd_rn(r.my, __dmul_rn(mb, (double)y[ch]));
            r.mz = __dadd_rn(r.mz, __dmul_rn(mb, (double)z[ch]));
            r.cnt += 1;
            r.first = min(r.first, pos[ch]);
        } else {
            r.ms = __dadd_rn(r.ms, mass64[ch]);
            r.mx = __dadd_rn(r.mx, mom64[3 * (int64_t)ch + 0]);
            r.my = __dadd_rn(r.my, mom64[3 * (int64_t)ch + 1]);
            r.mz = __dadd_rn(r.mz, mom64[3 * (int64_t)ch + 2]);
            r.cnt += count[ch];
            r.first = min(r.first, start[ch]);
        }
    }
    return r;
}

// Single system.  Thread t < C handles cell n + t, thread C the root (C = *index - n).
__global__ void __launch_bounds__(256) summarize_kernel(
    int n, const int *__restrict__ index, int *child, const float *x, const float *y, const float *z,
    float *mass, float *cx, float *cy, float *cz, int *count, int *start, const int32_t *__restrict__ pos,
    const int32_t *__restrict__ parent, int32_t *arrive, const int32_t *__restrict__ cell_level,
    double *mass64, double *mom64, float4 *nodes4, DeviceState *st)
{
    if (latched(st)) return;
    const int ncell = *index - n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ncell) return;
    int c = (t < ncell) ? n + t : 0;
    if (count_cell_children(child, c, n) != 0) return;
    for (;;) {
        const bool bucket = c != 0 && cell_level[c] == kKeyLevels;
        const CellSum r = sum_cell(c, bucket, n, child, x, y, z, mass, pos, mass64, mom64, count, start);
        if (!(r.ms > 0.0)) atomicOr(&st->err, kErrNonFinite);
        const double gx = r.mx / r.ms, gy = r.my / r.ms, gz = r.mz / r.ms;
        const float fx = (float)gx, fy = (float)gy, fz = (float)gz, fm = (float)r.ms;
        if (c == 0) {
            st->root64[0] = r.ms; st->root64[1] = gx; st->root64[2] = gy; st->root64[3] = gz;
            st->root4 = make_float4(fx, fy, fz, fm);
            break;
        }
        mass64[c] = r.ms;
        mom64[3 * (int64_t)c + 0] = r.mx; mom64[3 * (int64_t)c + 1] = r.my; mom64[3 * (int64_t)c + 2] = r.mz;
        cx[c] = fx; cy[c] = fy; cz[c] = fz; mass[c] = fm;
        count[c] = r.cnt;
        start[c] = r.first;
        nodes4[c] = make_float4(fx, fy, fz, fm);
        const int p = parent[c];
        __threadfence();
        const int old = atomicAdd(&arrive[p], 1);
        if (old != count_cell_children(child, p, n) - 1) break;
        __threadfence();
        c = p;
    }
}

// Ensemble.  Thread t handles cell n + t of [n, *index); root cells (parent -1)
// end the walk up.
__global__ void __launch_bounds__(256) summarize_batch_kernel(
    int n, const int *__restrict__ index, int *child, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ z, const float *__restrict__ mass, int *count, int *start,
    const int32_t *__restrict__ pos, const int32_t *__restrict__ parent, int32_t *arrive,
    const int32_t *__restrict__ cell_level, double *mass64, double *mom64, float4 *nodes4, DeviceState *st)
{
    if (latched(st)) return;
    const int ncell = *index - n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int c = n + t;
    if (count_cell_children(child, c, n) != 0) return;
    for (;;) {
        const CellSum r = sum_cell(c, cell_level[c] == kKeyLevels, n, child, x, y, z, mass, pos, mass64, mom64,
                                   count, start);
        if (r.cnt > 0 && !(r.ms > 0.0)) atomicOr(&st->err, kErrNonFinite);
        const double gx = r.mx / r.ms, gy = r.my / r.ms, gz = r.mz / r.ms;
        mass64[c] = r.ms;
        mom64[3 * (int64_t)c + 0] = r.mx; mom64[3 * (int64_t)c + 1] = r.my; mom64[3 * (int64_t)c + 2] = r.mz;
        count[c] = r.cnt;
        start[c] = r.first;
        nodes4[c] = make_float4((float)gx, (float)gy, (float)gz, (float)r.ms);
        const int p = parent[c];
        if (p < 0) break;
        __threadfence();
        const int old = atomicAdd(&arrive[p], 1);
        if (old != count_cell_children(child, p, n) - 1) break;
        __threadfence();
        c = p;
    }
}

// Traversal layout: bodies in key order as float4 (x, y, z, m) at nodes4[0..n),
// and child references with bodies renamed to their sorted positions
// (tchild[8c+s]); cells keep their ids.  Also writes sorted[].
__global__ void __launch_bounds__(256) layout_bodies_kernel(
    int n, const int32_t *__restrict__ perm, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ z, const float *__restrict__ mass, float4 *__restrict__ nodes4,
    int *__restrict__ sorted, const DeviceState *st)
{
    if (latched(st)) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) {
        const int b = perm[k];
        sorted[k] = b;
        nodes4[k] = make_float4(x[b], y[b], z[b], mass[b]);
    }
}

// Single system: item 0 is the root (child[0..7]), item i >= 1 is cell n + i - 1.
// Ensemble: item i is cell n + i (root cells included).
template <bool BATCH>
__global__ void __launch_bounds__(256) layout_child_kernel(
    int n, const int *__restrict__ index, const int *__restrict__ child, const int32_t *__restrict__ pos,
    int32_t *__restrict__ tchild, const DeviceState *st)
{
    if (latched(st)) return;
    const int ncell = *index - n;
    const int64_t items = BATCH ? ncell : ncell + 1;
    const int64_t total = 8 * items;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
         k += (int64_t)blockDim.x * gridDim.x) {
        const int64_t item = k >> 3;
        const int64_t node = BATCH ? n + item : (item == 0 ? 0 : n + item - 1);
        const int ch = child[8 * node + (k & 7)];
        tchild[8 * node + (k & 7)] = (ch >= 0 && ch < n) ? pos[ch] : ch;
    }
}

static inline int capped_grid(int64_t work, int cap = 65536)
{
    int64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

void launch_build(nbx_context *c, hipStream_t s, int n, int m, int *child, int *index)
{
    if (n <= 0) return;
    build_kernel<false><<<(n + 255) / 256, 256, 0, s>>>(n, m, c->skeys, c->perm, c->keys, child, index,
                                                        c->cell_level, c->parent, c->arrive, c->st, nullptr, 0);
}

void launch_summarize(nbx_context *c, hipStream_t s)
{
    const int threads = c->m - c->n + 1;
    summarize_kernel<<<(threads + 255) / 256, 256, 0, s>>>(
        c->n, c->index, c->child, c->x, c->y, c->z, c->mass, c->x, c->y, c->z, c->count, c->start, c->pos,
        c->parent, c->arrive, c->cell_level, c->mass64, c->mom64, c->nodes4, c->st);
}

void launch_layout(nbx_context *c, hipStream_t s, int *sorted)
{
    const int n = c->n;
    layout_bodies_kernel<<<capped_grid(n), 256, 0, s>>>(n, c->perm, c->x, c->y, c->z, c->mass, c->nodes4, sorted,
                                                        c->st);
    layout_child_kernel<false><<<capped_grid(8 * (int64_t)(c->m - c->n + 1)), 256, 0, s>>>(
        n, c->index, c->child, c->pos, c->tchild, c->st);
}

void launch_batch_build(nbx_batch *b, hipStream_t s)
{
    build_kernel<true><<<(b->N + 255) / 256, 256, 0, s>>>(b->N, b->m, b->skeys, b->perm, b->keys, b->child,
                                                          b->index, b->cell_level, b->parent, b->arrive, b->st,
                                                          b->sys_of, b->N);
}

void launch_batch_summarize(nbx_batch *b, hipStream_t s)
{
    const int threads = b->m - b->N;
    summarize_batch_kernel<<<(threads + 255) / 256, 256, 0, s>>>(
        b->N, b->index, b->child, b->x, b->y, b->z, b->mass, b->count, b->start, b->pos, b->parent, b->arrive,
        b->cell_level, b->mass64, b->mom64, b->nodes4, b->st);
}

void launch_batch_layout(nbx_batch *b, hipStream_t s)
{
    layout_bodies_kernel<<<capped_grid(b->N), 256, 0, s>>>(b->N, b->perm, b->x, b->y, b->z, b->mass, b->nodes4,
                                                           b->sorted, b->st);
    layout_child_kernel<true><<<capped_grid(8 * (int64_t)(b->m - b->N)), 256, 0, s>>>(
        b->N, b->index, b->child, b->pos, b->tchild, b->st);
}

}  // namespace nbx
