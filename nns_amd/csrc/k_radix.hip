// k_radix.hip -- stable LSD radix sort of (64-bit key, int32 index) pairs.
//
// Eight passes of 8 bits.  Each pass is three launches over tiles of 4096 keys
// (4 waves x 16 items x 64 lanes; wave w owns the contiguous quarter w of the
// tile, walked item by item so that lane order is element order):
//   upsweep   per-tile digit histogram           reads 8 B/key
//   scan      exclusive scan, digit-major          ~1 KiB per tile
//   downsweep stable rank + scatter                reads 12 B/key, writes 12 B/key
// Ranks inside a wave come from a 64-lane match of the digit built from eight
// ballots (no atomics, no intra-wave waits), so equal digits keep their order.
#include "nbx_internal.h"

namespace nbx {

__device__ inline uint64_t lanemask_lt(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

// Lanes of the wave whose `digit` equals this lane's; only lanes in `live` count.
__device__ inline uint64_t match_digit(uint32_t digit, uint64_t live)
{
    uint64_t peers = live;
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
        const bool bit = (digit >> b) & 1u;
        const uint64_t vote = __ballot(bit);
        peers &= bit ? vote : ~vote;
    }
    return peers;
}

__global__ void __launch_bounds__(kSortThreads) radix_upsweep(
    const uint64_t *__restrict__ keys, int n, int shift, int ntiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t cnt[kSortThreads / 64][kRadix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < (kSortThreads / 64) * kRadix; k += kSortThreads) (&cnt[0][0])[k] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)wave * (kSortItems * 64);
#pragma unroll 4
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        const bool valid = e < n;
        const uint32_t d = valid ? (uint32_t)((keys[e] >> shift) & (kRadix - 1)) : 0u;
        const uint64_t live = __ballot(valid);
        const uint64_t peers = match_digit(d, live);
        if (valid && lane == __ffsll((unsigned long long)peers) - 1) cnt[wave][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for (int d = tid; d < kRadix; d += kSortThreads) {
        uint32_t s = 0;
        for (int w = 0; w < kSortThreads / 64; ++w) s += cnt[w][d];
        hist[(int64_t)d * ntiles + blockIdx.x] = s;
    }
}

// Exclusive scan of `total` counters in one block of 1024 threads.
__global__ void __launch_bounds__(1024) radix_scan(uint32_t *__restrict__ hist, int total)
{
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (total + 1023) / 1024;
    const int lo = tid * per, hi = min(total, lo + per);
    uint32_t s = 0;
    for (int k = lo; k < hi; ++k) s += hist[k];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - s;
    for (int k = lo; k < hi; ++k) {
        uint32_t v = hist[k];
        hist[k] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kSortThreads) radix_downsweep(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin, int n, int shift, int ntiles,
    const uint32_t *__restrict__ hist, uint64_t *__restrict__ kout, int32_t *__restrict__ vout,
    int32_t *__restrict__ inv)
{
    constexpr int W = kSortThreads / 64;
    __shared__ uint32_t run[W][kRadix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < W * kRadix; k += kSortThreads) (&run[0][0])[k] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)wave * (kSortItems * 64);
    const uint64_t lt = lanemask_lt(lane);
    uint64_t key[kSortItems];
    int32_t val[kSortItems];
    uint32_t rank[kSortItems];
    uint32_t dig[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        const bool valid = e < n;
        key[i] = valid ? kin[e] : 0ull;
        val[i] = valid ? (vin ? vin[e] : (int32_t)e) : 0;
        const uint32_t d = (uint32_t)((key[i] >> shift) & (kRadix - 1));
        dig[i] = d;
        const uint64_t live = __ballot(valid);
        const uint64_t peers = match_digit(d, live);
        const uint32_t before = run[wave][d];
        rank[i] = before + (uint32_t)__popcll(peers & lt);
        if (valid && lane == __ffsll((unsigned long long)peers) - 1) run[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // exclusive prefix over the waves of this tile, plus the tile's global offset
    for (int d = tid; d < kRadix; d += kSortThreads) {
        uint32_t s = hist[(int64_t)d * ntiles + blockIdx.x];
        for (int w = 0; w < W; ++w) {
            uint32_t c = run[w][d];
            run[w][d] = s;
            s += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        if (e < n) {
            const uint32_t p = run[wave][dig[i]] + rank[i];
            kout[p] = key[i];
            vout[p] = val[i];
            if (inv) inv[val[i]] = (int32_t)p;
        }
    }
}

void launch_radix_passes(uint32_t *hist, hipStream_t s, const uint64_t *keys_in, const int32_t *vals_in, int n,
                         int first_shift, int npasses, uint64_t *keys_out, int32_t *vals_out, uint64_t *tmp_k,
                         int32_t *tmp_v, int32_t *inv_out)
{
    if (n <= 0 || npasses <= 0) return;
    const int ntiles = (n + kSortTile - 1) / kSortTile;
    const uint64_t *ki = keys_in;
    const int32_t *vi = vals_in;
    for (int p = 0; p < npasses; ++p) {
        const int shift = first_shift + p * kRadixBits;
        const bool last = (p == npasses - 1);
        // alternate so that the last pass lands in keys_out / vals_out
        const bool to_out = ((npasses - 1 - p) % 2) == 0;
        uint64_t *ko = to_out ? keys_out : tmp_k;
        int32_t *vo = to_out ? vals_out : tmp_v;
        radix_upsweep<<<ntiles, kSortThreads, 0, s>>>(ki, n, shift, ntiles, hist);
        radix_scan<<<1, 1024, 0, s>>>(hist, ntiles * kRadix);
        radix_downsweep<<<ntiles, kSortThreads, 0, s>>>(ki, vi, n, shift, ntiles, hist, ko, vo,
                                                        last ? inv_out : nullptr);
        ki = ko;
        vi = vo;
    }
}

// Batched form: after the 8 key passes, a stable pass over the system index
// groups every system's bodies, in key order, at [off[s], off[s+1]).
__global__ void __launch_bounds__(256) system_key_kernel(int n, const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ sys_of,
                                                         uint64_t *__restrict__ k2)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x)
        k2[k] = (uint64_t)sys_of[perm[k]];
}

__global__ void __launch_bounds__(256) gather_sorted_kernel(int n, const int32_t *__restrict__ perm,
                                                            const uint64_t *__restrict__ keys,
                                                            uint64_t *__restrict__ skeys,
                                                            int32_t *__restrict__ pos)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) {
        const int b = perm[k];
        skeys[k] = keys[b];
        pos[b] = k;
    }
}

void launch_batch_sort(nbx_batch *b, hipStream_t s)
{
    const int n = b->N;
    int g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    // 1. stable sort of (key, body) over the 64 key bits
    launch_radix_passes(b->hist, s, b->keys, nullptr, n, 0, 64 / kRadixBits, b->skeys, b->perm1, b->kbuf,
                        b->vbuf, nullptr);
    // 2. stable sort by system (16 bits cover 65536 systems)
    system_key_kernel<<<g, 256, 0, s>>>(n, b->perm1, b->sys_of, b->k2);
    launch_radix_passes(b->hist, s, b->k2, b->perm1, n, 0, 2, b->k2buf, b->perm, b->kbuf, b->vbuf, nullptr);
    // 3. keys in (system, key) order and the inverse permutation
    gather_sorted_kernel<<<g, 256, 0, s>>>(n, b->perm, b->keys, b->skeys, b->pos);
}

void launch_exclusive_scan_u32(uint32_t *data, int total, hipStream_t s)
{
    if (total > 0) radix_scan<<<1, 1024, 0, s>>>(data, total);
}

void launch_radix_sort(nbx_context *c, hipStream_t s, const uint64_t *keys_in, int n, uint64_t *keys_out,
                       int32_t *perm_out, uint64_t *tmp_k, int32_t *tmp_v, int32_t *inv_out)
{
    launch_radix_passes(c->hist, s, keys_in, nullptr, n, 0, 64 / kRadixBits, keys_out, perm_out, tmp_k, tmp_v,
                        inv_out);
}

}  // namespace nbx
