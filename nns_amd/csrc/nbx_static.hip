// nbx_static.hip -- nbx_sort_keys_static: the stable key sort of nbx_sort_keys
// for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops): no context, no hipMalloc, no hipFree.  Its scratch -- the
// ping-pong keys and values of the radix passes and the digit histograms -- is
// one file-scope device array, `scratch`, of kStaticSortMax keys' worth
// (12.5 MiB), which every GPU the library runs on holds until the process ends.
//
// Host code never passes `scratch` itself to a kernel: on the host the name is
// a stand-in, not a device address.  The first call on a device, with that
// device current, asks the runtime for the array's address there
// (hipGetSymbolAddress) and keeps it per device; every later call on that
// device uses the kept address.
//
// All calls on one GPU share that GPU's array, whose contents are those of the
// last call: calls that overlap on one GPU (two streams, two host threads) race.
// Nothing in the array needs to start at zero: each radix pass writes every
// histogram entry it later reads.
#include <mutex>

#include "nbx_internal.h"

namespace {

constexpr int kStaticSortMax = 1 << 20;  // keys per call
constexpr size_t kTmpKeyBytes = sizeof(uint64_t) * (size_t)kStaticSortMax;
constexpr size_t kTmpValBytes = sizeof(int32_t) * (size_t)kStaticSortMax;
constexpr size_t kHistBytes =
    sizeof(uint32_t) * ((size_t)(kStaticSortMax / nbx::kSortTile) * nbx::kRadix + nbx::kRadix);
constexpr size_t kScratchBytes = kTmpKeyBytes + kTmpValBytes + kHistBytes;
static_assert(kTmpKeyBytes % 256 == 0 && kTmpValBytes % 256 == 0, "sub-buffers stay 256-byte aligned");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once looked up

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

extern "C" int nbx_sort_keys_static(const uint64_t *keys, int n, uint64_t *sorted_keys, int32_t *perm,
                                    void *stream)
{
    if (n < 0 || n > kStaticSortMax) {
        nbx::set_create_error("nbx_sort_keys_static: n out of range [0, 1,048,576]");
        return NBX_ERR_INVALID;
    }
    if (!keys || !sorted_keys || !perm) {
        nbx::set_create_error("nbx_sort_keys_static: null array");
        return NBX_ERR_INVALID;
    }
    if ((const void *)keys == (void *)sorted_keys) {
        nbx::set_create_error("nbx_sort_keys_static: keys and sorted_keys alias");
        return NBX_ERR_INVALID;
    }
    if (n == 0) return NBX_OK;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        nbx::set_create_error(std::string("nbx_sort_keys_static: hipGetDevice: ") + hipGetErrorString(e));
        return NBX_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        nbx::set_create_error("nbx_sort_keys_static: device ordinal above 63");
        return NBX_ERR_INVALID;
    }
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!g_scratch[dev]) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                nbx::set_create_error(std::string("nbx_sort_keys_static: hipGetSymbolAddress: ") +
                                      hipGetErrorString(e));
                return NBX_ERR_HIP;
            }
            g_scratch[dev] = p;
        }
        base = static_cast<unsigned char *>(g_scratch[dev]);
    }
    uint64_t *tmp_k = reinterpret_cast<uint64_t *>(base);
    int32_t *tmp_v = reinterpret_cast<int32_t *>(base + kTmpKeyBytes);
    uint32_t *hist = reinterpret_cast<uint32_t *>(base + kTmpKeyBytes + kTmpValBytes);
    // 8 passes of 8 bits: the same launches as nbx_sort_keys over 64-bit keys
    nbx::launch_radix_passes<uint64_t>(hist, (hipStream_t)stream, keys, nullptr, n, 0, 64 / nbx::kRadixBits,
                                       sorted_keys, perm, tmp_k, tmp_v, nullptr);
    e = hipGetLastError();
    if (e != hipSuccess) {
        nbx::set_create_error(std::string("nbx_sort_keys_static: kernel launch: ") + hipGetErrorString(e));
        return NBX_ERR_HIP;
    }
    return NBX_OK;
}
