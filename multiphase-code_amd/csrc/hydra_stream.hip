// hydra_stream.hip -- the device's streaming copy rate, measured with a
// hand-written kernel: the ceiling beside which the hot-path kernels' HBM
// rates are read, and a kernel of known traffic (reads `bytes`, writes
// `bytes`) that calibrates the HBM byte counters.
#include <string>

#include "hydra_dev.h"
#include "hydra_hip.h"

namespace hydra {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int SC_BLOCK = 256;
constexpr int SC_UNROLL = 4;  // 16-byte words per thread, loads before stores

__global__ void __launch_bounds__(SC_BLOCK) k_stream_copy(const u32x4 *__restrict__ a, u32x4 *__restrict__ b,
                                                          size_t n)
{
    const size_t base = (size_t)blockIdx.x * (SC_BLOCK * SC_UNROLL) + threadIdx.x;
    u32x4 v[SC_UNROLL];
#pragma unroll
    for (int u = 0; u < SC_UNROLL; u++) {
        const size_t k = base + (size_t)u * SC_BLOCK;
        if (k < n) v[u] = __builtin_nontemporal_load(&a[k]);
    }
#pragma unroll
    for (int u = 0; u < SC_UNROLL; u++) {
        const size_t k = base + (size_t)u * SC_BLOCK;
        if (k < n) __builtin_nontemporal_store(v[u], &b[k]);
    }
}

}  // namespace hydra

extern "C" int hp_stream_bandwidth(size_t bytes, int reps, double *gbs, double *ms_per_copy)
{
    using namespace hydra;
    if (!gbs || reps <= 0 || bytes < 16 || bytes % 16)
        return set_global_error(HP_ERR_ARG, "hp_stream_bandwidth: bad argument");
    const size_t n = bytes / 16;
    u32x4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    const dim3 grid((unsigned)((n + SC_BLOCK * SC_UNROLL - 1) / (SC_BLOCK * SC_UNROLL)));
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipMemset(a, 1, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_stream_copy, grid, dim3(SC_BLOCK), 0, nullptr, a, b, n);  // warm-up
        e = hipEventRecord(e0, nullptr);
    }
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(k_stream_copy, grid, dim3(SC_BLOCK), 0, nullptr, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess)
        return set_global_error(HP_ERR_HIP, std::string("hp_stream_bandwidth: ") + hipGetErrorString(e));
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return HP_OK;
}
