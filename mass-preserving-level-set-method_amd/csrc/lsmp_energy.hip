// lsmp_energy.hip -- the per-step energy and momentum integrals
// (energy_and_momentum, Vortex/src/others/output.f90:250-397; the Disk/ and
// Vortex2D/ copies are the same text) with rho_mu (ns_solver/others.f90:47-80).
//
// The nine first derivatives of the velocity come from the mass correction's
// CCD line solves (launch_ccd_pass, every line of a direction; all three
// directions scale by 1 / dx as the reference's solvers do,
// ccd_solvers.f90:406-408).  One grid-stride pass then forms eleven block
// sums per workgroup with Neumaier-compensated FP64 accumulation, and one
// workgroup reduces the block sums in a fixed order:
//   [0] A dx    (cells with 0 < phi < dx, :322-324)  [1] v = sum marker dV
//   [2..4] sum marker (u, v, w) dV                   [5..7] sum marker (x, y, z) dV
//   [8] Ke, [9] Po, [10] Q over the cells with marker > 0.99 (:337-356).
// marker is H(phi) from ls_funs, or the VOF for the CLSVOF method (:326-330).
// In 2D the z derivatives of a field constant along z are zero, and w is
// zero (the 2D velocity never sets it).
#include <hip/hip_runtime.h>

#include "lsmp_device.h"
#include "lsmp_internal.h"

namespace lsmp {

namespace {

constexpr int ENV = ENERGY_SUMS;

template <typename T, int DIM, bool VOFM>
__global__ __launch_bounds__(256) void k_energy_sums(const T *__restrict__ phi, const T *__restrict__ vof,
                                                     const T *__restrict__ u, const T *__restrict__ v,
                                                     const T *__restrict__ w, const T *__restrict__ ten,
                                                     EnergyArgs ea, double *__restrict__ partial)
{
    double acc[ENV], cmp[ENV];
#pragma unroll
    for (int q = 0; q < ENV; ++q) acc[q] = cmp[q] = 0.0;
    const size_t n = (size_t)ea.nx * ea.ny * ea.nz;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % (size_t)ea.nx);
        const size_t r = q / (size_t)ea.nx;
        const int j = (int)(r % (size_t)ea.ny), k = (int)(r / (size_t)ea.ny);
        const double p = (double)phi[q];
        if (p < ea.dx && p > 0.0) csum_add(acc[0], cmp[0], ea.dv);
        double marker;
        if (VOFM) {
            marker = (double)vof[q];
        } else {
            double dd;
            ls_funs<double>(p, ea.ls_wid, marker, dd);
        }
        const double uu = (double)u[q], vv = (double)v[q], ww = DIM == 3 ? (double)w[q] : 0.0;
        // cell centres with (i - 1/2) in single precision (tree.f90:254-274)
        const double x = ea.xstart + (double)((float)(i + 1) - 0.5f) * ea.dxc;
        const double y = ea.ystart + (double)((float)(j + 1) - 0.5f) * ea.dyc;
        const double z = ea.zstart + (double)((float)(k + 1) - 0.5f) * ea.dzc;
        csum_add(acc[2], cmp[2], marker * uu * ea.dv);
        csum_add(acc[3], cmp[3], marker * vv * ea.dv);
        csum_add(acc[4], cmp[4], marker * ww * ea.dv);
        csum_add(acc[1], cmp[1], marker * ea.dv);
        csum_add(acc[5], cmp[5], marker * x * ea.dv);
        csum_add(acc[6], cmp[6], marker * y * ea.dv);
        csum_add(acc[7], cmp[7], marker * z * ea.dv);
        if (marker > 0.99) {
            double rho, mu;
            if (!ea.inverse) {
                rho = marker + ea.rho_12 * (1.0 - marker);
                mu = marker + ea.mu_12 * (1.0 - marker);
            } else {
                rho = marker * ea.rho_12 + (1.0 - marker);
                mu = marker * ea.mu_12 + (1.0 - marker);
            }
            const double lke = uu * uu + vv * vv + ww * ww;
            csum_add(acc[8], cmp[8], lke * rho * 0.5 * ea.dv);
            csum_add(acc[9], cmp[9], z * rho * ea.dv);
            // ten: d/dx_d of component f at ten[(3 f + d) n + q]
            auto d = [&](int f, int dir) { return (double)ten[(size_t)(3 * f + dir) * n + q]; };
            const double ux = d(0, 0), uy = d(0, 1), vx = d(1, 0), vy = d(1, 1);
            const double uz = DIM == 3 ? d(0, 2) : 0.0, vz = DIM == 3 ? d(1, 2) : 0.0;
            const double wx = DIM == 3 ? d(2, 0) : 0.0, wy = DIM == 3 ? d(2, 1) : 0.0, wz = DIM == 3 ? d(2, 2) : 0.0;
            const double lq = 0.5 * ((uy + vx) * (uy + vx) + (uz + wx) * (uz + wx) + (wy + vz) * (wy + vz)) +
                              ux * ux + vy * vy + wz * wz;
            csum_add(acc[10], cmp[10], lq * 2.0 * mu * ea.dv);
        }
    }
    block_sum_comp<ENV>(acc, cmp);
    if (threadIdx.x == 0)
        for (int q = 0; q < ENV; ++q) partial[(size_t)ENV * blockIdx.x + q] = acc[q];
}

// the block sums in a fixed order (one workgroup), then the ratios of :366-372
__global__ void k_energy_finalize(const double *__restrict__ partial, int nblk, double dx, double *__restrict__ out)
{
    double acc[ENV], cmp[ENV];
#pragma unroll
    for (int q = 0; q < ENV; ++q) acc[q] = cmp[q] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < ENV; ++q) csum_add(acc[q], cmp[q], partial[(size_t)ENV * b + q]);
    block_sum_comp<ENV>(acc, cmp);
    if (threadIdx.x == 0) {
        const double vol = acc[1];
        out[0] = acc[0] / dx;
        out[1] = vol;
        for (int q = 2; q < 8; ++q) out[q] = acc[q] / vol;
        out[8] = acc[8];
        out[9] = acc[9];
        out[10] = acc[10];
    }
}

}  // namespace

template <typename T>
void launch_energy_sums(hipStream_t st, int dim, bool vof_marker, const T *phi, const T *vof, const T *u, const T *v,
                        const T *w, const T *ten, const EnergyArgs &ea, double *partial, int nblocks)
{
#define LSMP_EN(D, V)                                                                                          \
    hipLaunchKernelGGL((k_energy_sums<T, D, V>), dim3(nblocks), dim3(256), 0, st, phi, vof, u, v, w, ten, ea,   \
                       partial)
    if (dim == 3) {
        if (vof_marker) LSMP_EN(3, true);
        else LSMP_EN(3, false);
    } else {
        if (vof_marker) LSMP_EN(2, true);
        else LSMP_EN(2, false);
    }
#undef LSMP_EN
}

void launch_energy_finalize(hipStream_t st, const double *partial, int nblk, double dx, double *out)
{
    hipLaunchKernelGGL(k_energy_finalize, dim3(1), dim3(1024), 0, st, partial, nblk, dx, out);
}

template void launch_energy_sums<double>(hipStream_t, int, bool, const double *, const double *, const double *,
                                         const double *, const double *, const double *, const EnergyArgs &,
                                         double *, int);
template void launch_energy_sums<float>(hipStream_t, int, bool, const float *, const float *, const float *,
                                        const float *, const float *, const float *, const EnergyArgs &, double *,
                                        int);

}  // namespace lsmp
