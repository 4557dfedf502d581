// lsmp_noalloc.hip -- lsmp_mass_vol_noalloc: the LS mass and volume sums of
// lsmp_mass_vol (ls_mv, ns_solver/others.f90:172-213) on a caller's device
// field, with no device allocation, no host synchronisation and no handle.
//
// The block partial sums live in `scratch`, a file-scope device array of
// LSMP_NOALLOC_PARTIALS triples (24 KiB) that each GPU holds from the first
// call on it until the process ends.  Its device address on the current GPU
// is looked up once per device (hipGetSymbolAddress) and kept; host code uses
// that address, never the array's name.  Every call on a GPU shares that
// GPU's array and leaves its contents behind, so calls that overlap on one GPU
// (two streams, two host threads) race on it.  The reduction's block count is
// capped at LSMP_NOALLOC_PARTIALS, so every grid fits the array; each call
// writes every slot it reads, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lsmp.h"
#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace {

using namespace lsmp;

constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[3 * LSMP_NOALLOC_PARTIALS];

double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the current GPU, looked up on the first call
// on that GPU.
int scratch_address(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return LSMP_E_HIP;
    if (dev < 0 || dev >= kMaxDevices) return LSMP_E_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LSMP_E_HIP;
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return LSMP_OK;
}

// The final reduction of k_finalize3 (1024 threads, the same compensated sum
// in the same order), writing (mass, vol) to the caller's buffer.
__global__ __launch_bounds__(1024) void k_mass_vol_out(const double *__restrict__ partial, int nblk,
                                                       double *__restrict__ out)
{
    double acc[3];
    sum_partials3(partial, nblk, acc);
    if (threadIdx.x == 0) {
        out[0] = acc[1];
        out[1] = acc[0];
    }
}

}  // namespace

extern "C" int lsmp_mass_vol_noalloc(const double *phi, int nx, int ny, int nz, int dim, double dx, double dy,
                                     double dz, double ls_wid, double rho_12, int inverse, double *out,
                                     void *stream)
{
    if (!phi || !out || nx < 1 || ny < 1 || nz < 1 || (dim != 2 && dim != 3) || (dim == 2 && nz != 1))
        return LSMP_E_INVALID;
    if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !(ls_wid > 0) || (inverse != 0 && inverse != 1))
        return LSMP_E_INVALID;
    const size_t n = (size_t)nx * ny * nz;
    // lsmp_mass_vol's block count: one block of 256 threads per 256 cells, at most 1024
    size_t nblk = (n + 255) / 256;
    if (nblk > 1024) nblk = 1024;
    if (nblk > (size_t)LSMP_NOALLOC_PARTIALS) return LSMP_E_INVALID;
    double *partial = nullptr;
    int rc = scratch_address(&partial);
    if (rc) return rc;
    CcdArgs ca;
    ca.dx = dx; ca.dy = dy; ca.dz = dz;
    ca.dv = dx * dy * dz;
    ca.ls_wid = ls_wid;
    ca.rho_12 = rho_12;
    ca.inverse = inverse;
    hipStream_t st = static_cast<hipStream_t>(stream);
    launch_mass_vol<double>(st, phi, n, ca, partial, (int)nblk);
    hipLaunchKernelGGL(k_mass_vol_out, dim3(1), dim3(1024), 0, st, partial, (int)nblk, out);
    return hipGetLastError() == hipSuccess ? LSMP_OK : LSMP_E_HIP;
}
