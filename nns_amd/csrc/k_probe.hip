// k_probe.hip -- issue-rate probes of the FP32 vector ALU (diagnostics).
//
// The force walk's arithmetic ceiling is set by FP32 VALU issue.  These kernels
// measure the two rates it is made of on the device at hand: v_fma_f32 (the scalar
// FP32 FMA; packed FP32 is disabled for all device code of this library) and
// v_rsq_f32.  Every lane runs `iters` rounds over 8 independent registers, so
// each round is 8 instructions with no dependence inside it; one float per
// lane goes to out so that nothing is left out by the compiler.  Both chains
// stay finite: x -> 0.999 x + 1e-3 tends to 1, x -> x^-1/2 tends to 1 from
// any positive start.
#include "nbx_internal.h"

namespace nbx {

constexpr int kProbeThreads = 256;

template <int KIND>
__global__ void __launch_bounds__(kProbeThreads) valu_probe_kernel(int iters, float *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = 1.0f + 1e-3f * (float)(t & 255) + 0.125f * (float)j;
    const float mul = 0.999f, add = 1e-3f;
#pragma unroll 4
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = KIND == 0 ? fmaf(a[j], mul, add) : __builtin_amdgcn_rsqf(a[j]);
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += a[j];
    out[t] = s;
}

}  // namespace nbx

extern "C" int nbx_probe_valu(int kind, int blocks, int iters, float *out, int64_t out_count, void *stream,
                              float *ms)
{
    if (kind < 0 || kind > 1 || blocks < 1 || blocks > (1 << 22) || iters < 1 || !out || !ms) return NBX_ERR_INVALID;
    if ((int64_t)blocks * nbx::kProbeThreads > out_count) return NBX_ERR_INVALID;  // one float per thread
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return NBX_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return NBX_ERR_HIP;
    }
    hipError_t e = hipEventRecord(e0, s);
    if (e == hipSuccess && kind == 0) nbx::valu_probe_kernel<0><<<blocks, nbx::kProbeThreads, 0, s>>>(iters, out);
    if (e == hipSuccess && kind == 1) nbx::valu_probe_kernel<1><<<blocks, nbx::kProbeThreads, 0, s>>>(iters, out);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e == hipSuccess ? NBX_OK : NBX_ERR_HIP;
}
