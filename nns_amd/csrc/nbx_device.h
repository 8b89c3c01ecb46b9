// nbx_device.h -- device helpers shared by the kernel files (internal linkage,
// one copy per code object).
#pragma once

#include "nbx_internal.h"

namespace nbx {

// ---------------------------------------------------------------- keys
// i = (uint)(MASK * (p - o) / maxSpan) in FP64, truncated
// (3D/OpenMP/OpenMP_NBody/src/Body.cpp:14-16).  maxSpan == 0 maps to 0.  A body
// inside the extent gives t in [0, MASK]; t is clamped to that range (NaN to 0)
// before the cast, so a body outside it cannot make a negative or over-range
// coordinate (whose conversion is undefined) -- values inside are unchanged.
static __device__ inline uint64_t quantise(float p, double o, double s)
{
    if (!(s > 0.0)) return 0;
    double t = kMaskCoord * ((double)p - o);
    t = t / s;
    t = t >= 0.0 ? t : 0.0;
    t = t <= kMaskCoord ? t : kMaskCoord;
    return (uint64_t)t;
}

// Spreads the low 21 bits of v to every third bit (bit j -> bit 3j).
static __device__ inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x001f00000000ffffull;
    v = (v | (v << 16)) & 0x001f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// Body::getKey: 2^63 + sum_j 8^j (x_j + 2 y_j + 4 z_j) (Body.cpp:19-31).
static __device__ inline uint64_t morton_key(float x, float y, float z, const double *ext)
{
    uint64_t ix = quantise(x, ext[0], ext[3]);
    uint64_t iy = quantise(y, ext[1], ext[3]);
    uint64_t iz = quantise(z, ext[2], ext[3]);
    return (1ull << 63) | spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
}

// Lebesgue2Hilbert state tables (MolecularDynamics/BarnesHutParallel/
// SpaceFillingCurves/src/Tree.cpp:71-80); the values define the key.
static __constant__ unsigned char c_dir[12][8] = {
    {8, 10, 3, 3, 4, 5, 4, 5}, {2, 2, 11, 9, 4, 5, 4, 5}, {7, 6, 7, 6, 8, 10, 1, 1},
    {7, 6, 7, 6, 0, 0, 11, 9}, {0, 8, 1, 11, 6, 8, 6, 11}, {10, 0, 9, 1, 10, 7, 9, 7},
    {10, 4, 9, 4, 10, 2, 9, 3}, {5, 8, 5, 11, 2, 8, 3, 11}, {4, 9, 0, 0, 7, 9, 2, 2},
    {1, 1, 8, 5, 3, 3, 8, 6}, {11, 5, 0, 0, 11, 6, 2, 2}, {1, 1, 4, 10, 3, 3, 7, 10}};
static __constant__ unsigned char c_hil[12][8] = {
    {0, 7, 3, 4, 1, 6, 2, 5}, {4, 3, 7, 0, 5, 2, 6, 1}, {6, 1, 5, 2, 7, 0, 4, 3},
    {2, 5, 1, 6, 3, 4, 0, 7}, {0, 1, 7, 6, 3, 2, 4, 5}, {6, 7, 1, 0, 5, 4, 2, 3},
    {2, 3, 5, 4, 1, 0, 6, 7}, {4, 5, 3, 2, 7, 6, 0, 1}, {0, 3, 1, 2, 7, 4, 6, 5},
    {2, 1, 3, 0, 5, 6, 4, 7}, {4, 7, 5, 6, 3, 0, 2, 1}, {6, 5, 7, 4, 1, 2, 0, 3}};

// The state machine as one 96-byte table for a copy in shared memory: entry
// 8 dir + cell holds 8 c_dir[dir][cell] + c_hil[dir][cell], i.e. the next
// state's row offset and the digit.  Indexed per lane, the constant tables make
// every level a dependent divergent memory read; from shared memory a level
// costs one LDS read.  A kernel fills its copy with hilbert_table_fill, then
// waits at a workgroup barrier before the first lookup.
constexpr int kHilbertTable = 96;
static __device__ inline void hilbert_table_fill(uint8_t *s_tab)
{
    for (int t = threadIdx.x; t < kHilbertTable; t += blockDim.x)
        s_tab[t] = (uint8_t)(8 * c_dir[t >> 3][t & 7] + c_hil[t >> 3][t & 7]);
}

// The first `levels` Hilbert digits of a Lebesgue key (Lebesgue2Hilbert,
// SFC/src/Tree.cpp:82-104; levels = 21 maps all of them).  The mapping runs
// from the top level down, so stopping early leaves the coarse digits exact.
static __device__ inline uint64_t hilbert_top_lds(uint64_t lebesgue, int levels, const uint8_t *s_tab)
{
    uint64_t h = 0;
    int row = 0;
    for (int lvl = kKeyLevels; lvl > kKeyLevels - levels; --lvl) {
        const int e = s_tab[row + (int)((lebesgue >> ((lvl - 1) * 3)) & 7u)];
        h = (h << 3) + (uint64_t)(e & 7);
        row = e & ~7;
    }
    return h;
}

// Block-wide exclusive scan of one value per thread (blockDim.x a multiple of 64,
// at most 1024); returns the thread's exclusive prefix and writes the block total
// to *total.  Each wave scans its 64 values with six shuffles, the wave totals
// meet in sh[0 .. blockDim.x / 64): two barriers in all.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        const uint32_t t = sh[w];
        pre += w < wave ? t : 0u;
        tot += t;
    }
    *total = tot;
    __syncthreads();  // sh is free again for the caller
    return pre + x - v;
}

// First record of frame block b in the packed layout: the sum within its run of
// kOffRun blocks plus the records of the runs before it.
__device__ inline int frame_off(const uint32_t *__restrict__ boff, const uint32_t *__restrict__ bsum, int b)
{
    return (int)(boff[b] + bsum[b / kOffRun]);
}

// ---------------------------------------------------------------- tree
static __device__ inline int digit_at(uint64_t key, int level) { return (int)((key >> (60 - 3 * level)) & 7u); }

static __device__ inline int ld_slot(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ inline void st_slot(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ inline int cas_slot(int *p, int expect, int desired)
{
    __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expect;  // value found (== the expected value on success)
}

static __device__ inline int latched(const DeviceState *st)
{
    return __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- launch shape
// Hardware dispatches consecutive blocks round-robin over the 8 XCDs; this maps
// physical block b to a logical block so that each XCD gets a contiguous range.
static __device__ inline int xcd_swizzle(int b, int nblk)
{
    const int q = nblk / 8, r = nblk % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

static __device__ inline void kahan_add(float &sum, float &comp, float v)
{
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

}  // namespace nbx
