// k_radix.hip -- stable LSD radix sort of (64- or 32-bit key, int32 index) pairs.
//
// Passes of kRadixBits = 8 bits.  Each pass is three launches over tiles of 2048
// keys (4 waves x 8 items x 64 lanes; wave w owns the contiguous quarter w of the
// tile, walked item by item so that lane order is element order):
//   upsweep   per-tile digit histogram (LDS atomics) reads 8 B/key (64-bit keys)
//   scan      exclusive scan of each digit row       ~1 KiB per tile
//   downsweep stable rank, the tile reordered by digit in LDS, then written out
//             digit run by digit run (coalesced)  reads 12 B/key, writes 12 B/key
// (32-bit keys: 4 and 8 B/key).
// In the downsweep, ranks inside a wave come from a 64-lane match of the digit
// built from kRadixBits ballots (no atomics, no intra-wave waits), so equal
// digits keep their order; the upsweep needs counts only and takes them with
// LDS atomics.
#include "nbx_device.h"

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

template <class K>
__global__ void __launch_bounds__(kSortThreads) radix_upsweep(
    const K *__restrict__ keys, int n, int shift, int ntiles, uint32_t *__restrict__ hist,
    const int *__restrict__ n_dev)
{
    __shared__ uint32_t cnt[kSortThreads / 64][kRadix];
    if (n_dev) n = min(n, *n_dev);  // keys beyond the device-side count take no part
    if ((int64_t)blockIdx.x * kSortTile >= n) return;  // a tile past the count: not scanned, not read
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < (kSortThreads / 64) * kRadix; k += kSortThreads) (&cnt[0][0])[k] = 0u;
    __syncthreads();
    // Counts only (no ranks): one LDS atomic per key into the wave's own row;
    // lanes of a wave that share a digit are serialised by the LDS, which is
    // never slower than building the 64-lane digit match out of ballots.
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)wave * (kSortItems * 64);
    uint32_t dig[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        dig[i] = e < n ? (uint32_t)((keys[e] >> shift) & (kRadix - 1)) : kRadix;  // kRadix: no key
    }
#pragma unroll
    for (int i = 0; i < kSortItems; ++i)
        if (dig[i] < kRadix) atomicAdd(&cnt[wave][dig[i]], 1u);
    __syncthreads();
    for (int d = tid; d < kRadix; d += kSortThreads) {
        uint32_t s = 0;
        for (int w = 0; w < kSortThreads / 64; ++w) s += cnt[w][d];
        hist[(int64_t)d * ntiles + blockIdx.x] = s;
    }
}

// Exclusive scan of every digit row hist[d][0..ntiles) in place (one block per
// digit); row totals go to rowsum[d].  The digit bases (exclusive scan of the
// row totals) are formed by each downsweep block.  With a device-side count only
// the tiles that hold keys are scanned.
__global__ void __launch_bounds__(256) radix_rows_scan(uint32_t *__restrict__ hist, int ntiles,
                                                       uint32_t *__restrict__ rowsum, int n,
                                                       const int *__restrict__ n_dev)
{
    __shared__ uint32_t sh[256];
    uint32_t *row = hist + (int64_t)blockIdx.x * ntiles;
    const int used = n_dev ? (int)min((int64_t)ntiles, ((int64_t)min(n, *n_dev) + kSortTile - 1) / kSortTile) : ntiles;
    const int per = (used + 255) / 256;
    const int lo = threadIdx.x * per, hi = min(used, lo + per);
    uint32_t s = 0;
    for (int k = lo; k < hi; ++k) s += row[k];
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, sh, &total);
    for (int k = lo; k < hi; ++k) {
        const uint32_t v = row[k];
        row[k] = run;
        run += v;
    }
    if (threadIdx.x == 0) rowsum[blockIdx.x] = total;
}

// Generic exclusive scan of `total` uint32 (<= 1024 * 1024): per-block scans of
// 1024 elements, a scan of the block sums, then the block offsets added.
__global__ void __launch_bounds__(256) scan_blocks_kernel(uint32_t *__restrict__ data, int total,
                                                          uint32_t *__restrict__ sums)
{
    __shared__ uint32_t sh[256];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    uint32_t v[4], s = 0;
    for (int j = 0; j < 4; ++j) {
        v[j] = base + j < total ? data[base + j] : 0u;
        s += v[j];
    }
    uint32_t blk;
    uint32_t run = block_exclusive_scan(s, sh, &blk);
    for (int j = 0; j < 4; ++j) {
        if (base + j < total) data[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = blk;
}

__global__ void __launch_bounds__(1024) scan_sums_kernel(uint32_t *__restrict__ sums, int nb)
{
    __shared__ uint32_t sh[1024];
    const uint32_t v = threadIdx.x < nb ? sums[threadIdx.x] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, sh, &tot);
    if (threadIdx.x < nb) sums[threadIdx.x] = ex;
}

__global__ void __launch_bounds__(256) scan_add_kernel(uint32_t *__restrict__ data, int total,
                                                       const uint32_t *__restrict__ sums)
{
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    const uint32_t off = sums[blockIdx.x];
    for (int j = 0; j < 4; ++j)
        if (base + j < total) data[base + j] += off;
}

template <class K>
__global__ void __launch_bounds__(kSortThreads) radix_downsweep(
    const K *__restrict__ kin, const int32_t *__restrict__ vin, int n, int shift, int ntiles,
    const uint32_t *__restrict__ hist, const uint32_t *__restrict__ rowsum, K *__restrict__ kout,
    int32_t *__restrict__ vout, int32_t *__restrict__ inv, const int *__restrict__ n_dev)
{
    if (n_dev) n = min(n, *n_dev);
    if ((int64_t)blockIdx.x * kSortTile >= n) return;  // no keys in this tile
    constexpr int W = kSortThreads / 64;
    __shared__ uint32_t run[W][kRadix];
    __shared__ uint32_t dbase[kRadix];
    __shared__ uint32_t gdelta[kRadix];
    __shared__ uint32_t scan_sh[kSortThreads];
    __shared__ K s_key[kSortTile];
    __shared__ int32_t s_val[kSortTile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < W * kRadix; k += kSortThreads) (&run[0][0])[k] = 0u;
    {   // digit bases: exclusive scan of the digit totals, kRadix / kSortThreads digits per thread
        constexpr int R = kRadix / kSortThreads;
        static_assert(R * kSortThreads == kRadix, "digits per thread");
        uint32_t v[R], sum = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = rowsum[R * tid + r];
            sum += v[r];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(sum, dbase, &tot);  // dbase is the scan's scratch
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dbase[R * tid + r] = ex;
            ex += v[r];
        }
    }
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)wave * (kSortItems * 64);
    const uint64_t lt = lanemask_lt(lane);
    K key[kSortItems];
    int32_t val[kSortItems];
    uint32_t rank[kSortItems];
    uint32_t dig[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        const bool valid = e < n;
        key[i] = valid ? kin[e] : K(0);
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
    // The tile in digit order: digit d starts at lstart[d] = the tile's elements of
    // smaller digits, wave w's elements of digit d follow those of waves < w.  Each
    // digit's elements go to the global positions from gstart[d] (the tiles and
    // digits before), so global = local + (gstart[d] - lstart[d]).
    {
        constexpr int R = kRadix / kSortThreads;
        uint32_t cnt[R][W], tc[R], tsum = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int d = R * tid + r;
            tc[r] = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                cnt[r][w] = run[w][d];
                tc[r] += cnt[r][w];
            }
            tsum += tc[r];
        }
        uint32_t tot;
        uint32_t ls = block_exclusive_scan(tsum, scan_sh, &tot);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int d = R * tid + r;
            gdelta[d] = hist[(int64_t)d * ntiles + blockIdx.x] + dbase[d] - ls;
            uint32_t o = ls;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                run[w][d] = o;
                o += cnt[r][w];
            }
            ls += tc[r];
        }
    }
    __syncthreads();
    // scatter into the tile's digit order in LDS ...
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        const int64_t e = base + i * 64 + lane;
        if (e < n) {
            const uint32_t lp = run[wave][dig[i]] + rank[i];
            s_key[lp] = key[i];
            s_val[lp] = val[i];
        }
    }
    __syncthreads();
    // ... and write it out in that order: consecutive threads, consecutive addresses
    const int64_t t0 = (int64_t)blockIdx.x * kSortTile;
    const int tn = (int)min((int64_t)kSortTile, (int64_t)n - t0);
    for (int j = tid; j < tn; j += kSortThreads) {
        const K k = s_key[j];
        const int32_t v = s_val[j];
        const uint32_t d = (uint32_t)((k >> shift) & (kRadix - 1));
        const uint32_t p = (uint32_t)j + gdelta[d];
        kout[p] = k;
        vout[p] = v;
        if (inv) inv[v] = (int32_t)p;
    }
}

template <class K>
void launch_radix_passes(uint32_t *hist, hipStream_t s, const K *keys_in, const int32_t *vals_in, int n,
                         int first_shift, int npasses, K *keys_out, int32_t *vals_out, K *tmp_k,
                         int32_t *tmp_v, int32_t *inv_out, const int *n_dev)
{
    if (n <= 0 || npasses <= 0) return;
    const int ntiles = (n + kSortTile - 1) / kSortTile;
    const K *ki = keys_in;
    const int32_t *vi = vals_in;
    for (int p = 0; p < npasses; ++p) {
        const int shift = first_shift + p * kRadixBits;
        const bool last = (p == npasses - 1);
        // alternate so that the last pass lands in keys_out / vals_out
        const bool to_out = ((npasses - 1 - p) % 2) == 0;
        K *ko = to_out ? keys_out : tmp_k;
        int32_t *vo = to_out ? vals_out : tmp_v;
        uint32_t *rowsum = hist + (int64_t)ntiles * kRadix;  // kRadix totals after the rows
        radix_upsweep<K><<<ntiles, kSortThreads, 0, s>>>(ki, n, shift, ntiles, hist, n_dev);
        radix_rows_scan<<<kRadix, 256, 0, s>>>(hist, ntiles, rowsum, n, n_dev);
        radix_downsweep<K><<<ntiles, kSortThreads, 0, s>>>(ki, vi, n, shift, ntiles, hist, rowsum, ko, vo,
                                                        last ? inv_out : nullptr, n_dev);
        ki = ko;
        vi = vo;
    }
}

template void launch_radix_passes<uint64_t>(uint32_t *, hipStream_t, const uint64_t *, const int32_t *, int, int, int,
                                            uint64_t *, int32_t *, uint64_t *, int32_t *, int32_t *, const int *);
template void launch_radix_passes<uint32_t>(uint32_t *, hipStream_t, const uint32_t *, const int32_t *, int, int, int,
                                            uint32_t *, int32_t *, uint32_t *, int32_t *, int32_t *, const int *);

// Batched form: after the 8 key passes, stable passes over the system index
// (32-bit keys, one pass per 8 bits of it) group every system's bodies, in key
// order, at [off[s], off[s+1]).
__global__ void __launch_bounds__(256) system_key_kernel(int n, const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ sys_of,
                                                         uint32_t *__restrict__ k2)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x)
        k2[k] = (uint32_t)sys_of[perm[k]];
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
    // 1. stable sort of (key, body) over the 63 key bits below the constant top bit
    launch_radix_passes(b->hist, s, b->keys, nullptr, n, 0, (kMortonSortBits + kRadixBits - 1) / kRadixBits,
                        b->skeys, b->perm1, b->kbuf, b->vbuf, nullptr);
    // 2. stable sort by system: 32-bit keys, one pass up to 256 systems, two up to 65536
    int sys_bits = 0;
    while ((1 << sys_bits) < b->S) ++sys_bits;
    const int sys_passes = sys_bits > kRadixBits ? 2 : 1;
    uint32_t *k2 = reinterpret_cast<uint32_t *>(b->k2);
    system_key_kernel<<<g, 256, 0, s>>>(n, b->perm1, b->sys_of, k2);
    launch_radix_passes(b->hist, s, k2, b->perm1, n, 0, sys_passes, reinterpret_cast<uint32_t *>(b->k2buf), b->perm,
                        reinterpret_cast<uint32_t *>(b->kbuf), b->vbuf, nullptr);
    // 3. keys in (system, key) order and the inverse permutation
    gather_sorted_kernel<<<g, 256, 0, s>>>(n, b->perm, b->keys, b->skeys, b->pos);
}

// Exclusive scan of `total` uint32 in place by one workgroup of 1024 threads,
// 16 entries per thread per pass, the passes chained by a carry: one launch,
// for arrays of a few thousand entries.
__global__ void __launch_bounds__(1024) scan_one_block_kernel(uint32_t *__restrict__ data, int total)
{
    __shared__ uint32_t sh[16];
    constexpr int V = 16;
    uint32_t carry = 0;
    for (int t0 = 0; t0 < total; t0 += V * (int)blockDim.x) {
        const int i0 = t0 + V * (int)threadIdx.x;
        uint32_t v[V], sum = 0;
#pragma unroll
        for (int u = 0; u < V; ++u) {
            v[u] = i0 + u < total ? data[i0 + u] : 0u;
            sum += v[u];
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(sum, sh, &tot) + carry;
#pragma unroll
        for (int u = 0; u < V; ++u) {
            if (i0 + u < total) data[i0 + u] = ex;
            ex += v[u];
        }
        carry += tot;
    }
}

void launch_scan_one_block(uint32_t *data, int total, hipStream_t s)
{
    if (total > 0) scan_one_block_kernel<<<1, 1024, 0, s>>>(data, total);
}

void launch_exclusive_scan_u32(uint32_t *data, int total, uint32_t *sums, hipStream_t s)
{
    if (total <= 0) return;
    const int nb = (total + 1023) / 1024;  // <= 1024 blocks (total <= 1M)
    scan_blocks_kernel<<<nb, 256, 0, s>>>(data, total, sums);
    scan_sums_kernel<<<1, 1024, 0, s>>>(sums, nb);
    scan_add_kernel<<<nb, 256, 0, s>>>(data, total, sums);
}

void launch_radix_sort(nbx_context *c, hipStream_t s, const uint64_t *keys_in, int n, uint64_t *keys_out,
                       int32_t *perm_out, uint64_t *tmp_k, int32_t *tmp_v, int32_t *inv_out, int key_bits)
{
    launch_radix_passes(c->hist, s, keys_in, nullptr, n, 0, (key_bits + kRadixBits - 1) / kRadixBits, keys_out,
                        perm_out, tmp_k, tmp_v, inv_out);
}

}  // namespace nbx
