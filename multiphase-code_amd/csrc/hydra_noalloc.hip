// hydra_noalloc.hip -- an entry point that allocates no device memory.
//
// hp_dt_min_noalloc is the CFL min-reduction of hp_dt_min (main.F90:520-548)
// over device arrays the caller owns, for callers that cannot allocate (code
// under HIP graph capture, serving loops that forbid allocation).  Its one
// piece of scratch memory, a partial minimum per block of cells, lives in a
// file-scope device array of fixed size that each GPU holds from the first
// use until the process ends.  Host code never passes the array's name to a
// kernel: the name is a host-side stand-in, so the device address is looked
// up once per device with hipGetSymbolAddress and kept.
#include <mutex>
#include <string>

#include "hydra_dev.h"
#include "hydra_hip.h"

namespace hydra {

constexpr int NA_BLOCK = 256;                                    // cells per partial
constexpr int NA_MAX_PARTS = HP_NOALLOC_MAX_CELLS / NA_BLOCK;    // 262,144 partials, 2 MiB
static_assert(NA_MAX_PARTS * NA_BLOCK == HP_NOALLOC_MAX_CELLS, "max cells is a multiple of the block");

static __device__ __align__(256) double scratch[NA_MAX_PARTS];

constexpr int NA_MAX_DEVICES = 64;
static double *g_scratch[NA_MAX_DEVICES];  // device address of scratch, per device ordinal
static std::mutex g_scratch_mu;

// Resolve the scratch address on the current device once; later calls read it.
static int scratch_of_current(double **out, std::string *err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *err = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return HP_ERR_HIP;
    }
    if (dev < 0 || dev >= NA_MAX_DEVICES) {
        *err = "device ordinal beyond the no-allocation table";
        return HP_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return HP_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return HP_OK;
}

int noalloc_setup(std::string *err)
{
    double *p = nullptr;
    return scratch_of_current(&p, err);
}

__device__ __forceinline__ double na_min(double a, double b) { return a < b ? a : b; }

// One partial per block: the minimum CFL bound of its NA_BLOCK cells
__global__ void __launch_bounds__(NA_BLOCK)
k_dt_part_noalloc(const double *q, size_t ldq, const double *dl, int n, double gamag, double gamap,
                  double pinf, double *part)
{
    const int i = blockIdx.x * NA_BLOCK + (int)threadIdx.x;
    double v = 1.0e8;
    if (i < n) v = cfl_bound(q, ldq, dl, i, gamag, gamap, pinf);
    for (int off = 32; off > 0; off >>= 1) v = na_min(v, __shfl_xor(v, off, 64));
    __shared__ double red[NA_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < NA_BLOCK / 64; w++) r = na_min(r, red[w]);
        part[blockIdx.x] = r;
    }
}

// The minimum of the nparts partials into out[0]; without a NaN bound a
// minimum does not depend on the order it is taken in, so this equals
// hp_dt_min's host reduction (with NaN bounds, a state that has broken
// down, hp_dt_min follows the reference's cell order and this does not)
__global__ void __launch_bounds__(1024) k_dt_final_noalloc(const double *part, int nparts, double *out)
{
    double v = 1.0e8;
    for (int t = (int)threadIdx.x; t < nparts; t += 1024) v = na_min(v, part[t]);
    for (int off = 32; off > 0; off >>= 1) v = na_min(v, __shfl_xor(v, off, 64));
    __shared__ double red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < 16; w++) r = na_min(r, red[w]);
        out[0] = r;
    }
}

}  // namespace hydra

extern "C" int hp_dt_min_noalloc(const double *q, size_t ldq, const double *dl, int neles, double gamag,
                                 double gamap, double pinf, double *dtmin, void *stream)
{
    using namespace hydra;
    if (!q || !dl || !dtmin || neles <= 0 || ldq < (size_t)neles)
        return set_global_error(HP_ERR_ARG, "hp_dt_min_noalloc: bad argument");
    if (neles > HP_NOALLOC_MAX_CELLS)
        return set_global_error(HP_ERR_ARG, "hp_dt_min_noalloc: neles exceeds HP_NOALLOC_MAX_CELLS (" +
                                                std::to_string(HP_NOALLOC_MAX_CELLS) + ")");
    hipStream_t st = static_cast<hipStream_t>(stream);
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return set_global_error(HP_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (st) {
        int sdev = -1;
        e = hipStreamGetDevice(st, &sdev);
        if (e != hipSuccess)
            return set_global_error(HP_ERR_HIP, std::string("hipStreamGetDevice: ") + hipGetErrorString(e));
        if (sdev != cur)
            return set_global_error(HP_ERR_ARG, "hp_dt_min_noalloc: the stream belongs to another device "
                                                "than the current one");
    }
    double *part = nullptr;
    std::string err;
    if (int rc = scratch_of_current(&part, &err)) return set_global_error(rc, "hp_dt_min_noalloc: " + err);
    const int nparts = (neles + NA_BLOCK - 1) / NA_BLOCK;
    hipLaunchKernelGGL(k_dt_part_noalloc, dim3(nparts), dim3(NA_BLOCK), 0, st, q, ldq, dl, neles, gamag, gamap,
                       pinf, part);
    hipLaunchKernelGGL(k_dt_final_noalloc, dim3(1), dim3(1024), 0, st, part, nparts, dtmin);
    e = hipGetLastError();
    if (e != hipSuccess) return set_global_error(HP_ERR_HIP, std::string("hp_dt_min_noalloc: ") + hipGetErrorString(e));
    return HP_OK;
}
